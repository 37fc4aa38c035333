/*
 * competitors.h -- CPU restatement of the baseline directory's competitor
 * tables (test infrastructure; see competitors.c).
 */
#ifndef ORACLE_COMPETITORS_H
#define ORACLE_COMPETITORS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_COMP_LINEAR = 0, ORC_COMP_CUCKOO = 1, ORC_COMP_HOPSCOTCH = 2, ORC_COMP_ROARING = 3 };

#define ORC_CUCKOO_STASH 16   /* cuckoo_hash.cuh:47 STASH_SIZE */
#define ORC_CUCKOO_KICKS 6000 /* cuckoo_hash.cuh:46 MAX_ITERATIONS */
#define ORC_HOP_H 32          /* hopscotch_hash.cuh:45 H */

/* One built table of any kind (arrays not used by a kind may be NULL). */
typedef struct {
  int32_t kind;
  const int64_t* toff;         /* linear / cuckoo / hopscotch: region offsets [n+1] */
  const int32_t* slots;        /* their slot arrays */
  const int32_t* stash_count;  /* cuckoo [n] */
  const int32_t* stash;        /* cuckoo [16 n] */
  const uint32_t* bitmap;      /* hopscotch, one word per slot */
  const int32_t* remap;        /* roaring [remap_len] */
  int64_t remap_len;
  const int32_t* node_start;   /* roaring [n+1] */
  const int32_t* word_index;
  const uint64_t* word_data;
} orc_comp;

uint32_t orc_fmix32(uint32_t h);
int64_t orc_comp_linear_geometry(const int32_t* off, int64_t n, float lf, int64_t* toff);
int64_t orc_comp_linear_build(const int32_t* off, const int32_t* el, int64_t n, const int64_t* toff,
                              int32_t* slots);
int64_t orc_comp_linear_check(const int32_t* off, const int32_t* el, int64_t n,
                              const int64_t* toff, const int32_t* slots);
int64_t orc_comp_cuckoo_geometry(const int32_t* off, int64_t n, float lf, int64_t* toff);
int64_t orc_comp_cuckoo_build(const int32_t* off, const int32_t* el, int64_t n,
                              const int64_t* toff, int32_t* table, int32_t* stash_count,
                              int32_t* stash);
int64_t orc_comp_hopscotch_geometry(const int32_t* off, int64_t n, float lf, int64_t* toff);
int64_t orc_comp_hopscotch_build(const int32_t* off, const int32_t* el, int64_t n,
                                 const int64_t* toff, int32_t* table, uint32_t* bitmap);
int64_t orc_comp_roaring_remap(const int32_t* off, const int32_t* el, int64_t n, int32_t* remap,
                               int64_t remap_len);
int64_t orc_comp_roaring_words(const int32_t* off, const int32_t* el, int64_t n,
                               const int32_t* remap);
int64_t orc_comp_roaring_build(const int32_t* off, const int32_t* el, int64_t n,
                               const int32_t* remap, int32_t* node_start, int32_t* word_index,
                               uint64_t* word_data);
int orc_comp_contains(const orc_comp* c, int32_t set, int32_t key);
uint64_t orc_comp_tc(const orc_comp* c, const int32_t* off, const int32_t* cols, int64_t n);
void orc_comp_sss(const orc_comp* c, const int32_t* off, const int32_t* el, const int32_t* cu,
                  const int32_t* cv, int64_t P, int32_t p, int32_t q, int32_t* flags,
                  int32_t* sizes);

#ifdef __cplusplus
}
#endif

#endif
