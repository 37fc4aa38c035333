/*
 * reset_amd_standalone.hpp -- the standalone applications' free-function
 * boundary, over the C ABI.
 *
 * The reference's triangle-count, set-similarity-search and
 * information-retrieval apps build their tables with
 *     void make_hash_table(float load_factor, int bucket_size);
 * which reads the app's CSR from global vectors and publishes the tables
 * through global device pointers (triangle-count/TC.cuh:18-38,
 * set-similarity-search/sss.cuh:22-43, information-retrieval/ir.cuh:19-33,
 * defined in each app's hashtable.cu).  An app keeps its own main (which
 * defines the input globals, as the reference mains do) and links, instead of
 * its hashtable.cu, one of the objects built from
 * reset_amd/csrc/standalone/make_hash_table.cpp:
 *     reset_amd/standalone/tc_hashtable.o   (compiled with RESET_AMD_APP_TC)
 *     reset_amd/standalone/sss_hashtable.o  (RESET_AMD_APP_SSS)
 *     reset_amd/standalone/ir_hashtable.o   (RESET_AMD_APP_IR)
 * plus libreset_amd.so.  Define the same macro before including this header.
 *
 * The globals keep the reference's names and types; only the ones an app's
 * hashtable.cu defines are defined by the object (the d_* outputs,
 * max_length, bucket_num, and the host mirrors hash_length / hash_start and
 * hash_table_* kept for link compatibility).  Differences: errors throw
 * reset_amd::Error instead of printing; bucket_num is checked to fit the app's
 * type (the TC app declares it int, which overflows at R-MAT scale 24); the
 * tables are owned by a handle (reset_amd_standalone_tables()) that
 * free_hash_table() releases, where the reference leaks them.
 */
#ifndef RESET_AMD_STANDALONE_HPP
#define RESET_AMD_STANDALONE_HPP

#include <vector>

#include "reset_amd.h"

#if defined(RESET_AMD_APP_TC)
extern std::vector<int> csr_row;   /* inputs, defined by the app's main */
extern std::vector<int> csr_cols;
extern int num_nodes;
extern int num_edges;
extern int* d_hash_table_hierarchical; /* outputs, defined by the object */
extern int* d_hash_table_normal;
extern int *d_hash_length, *d_csr_offsets;
extern long long* d_hash_tables_offset;
extern int max_length;
extern int bucket_num;
extern int* hash_table_normal;
extern int* hash_table_hierarchical;
extern std::vector<int> hash_length;
extern std::vector<int> hash_start;
extern int* d_csr_cols;
extern int* d_csr_row;
#elif defined(RESET_AMD_APP_SSS)
extern std::vector<int> vertexs;   /* inputs */
extern std::vector<int> csr_cols;
extern std::vector<int> csr_offsets;
extern int num_nodes;
extern int num_edges;
extern int* d_hash_table_hierarchical; /* outputs */
extern int* d_hash_table_normal;
extern int *d_hash_length, *d_csr_offsets;
extern long long* d_hash_tables_offset;
extern int max_length;
extern long long bucket_num;
extern int* hash_table_normal;
extern int* hash_table_hierarchical;
extern std::vector<int> hash_length;
extern std::vector<int> hash_start;
extern int* d_csr_cols;
extern int* d_vertexs;
#elif defined(RESET_AMD_APP_IR)
extern std::vector<int> inverted_index_offsets; /* inputs */
extern std::vector<int> inverted_index;
extern int inverted_index_num;
extern int* d_hash_table_hierarchical; /* outputs */
extern int* d_hash_table_normal;
extern int *d_hash_length, *d_inverted_index_offsets;
extern long long* d_hash_tables_offset;
extern int max_length;
extern long long bucket_num;
extern int* hash_table_normal;
extern int* hash_table_hierarchical;
extern std::vector<int> hash_length;
extern std::vector<int> hash_start;
extern int* d_inverted_index;
#else
#error "define RESET_AMD_APP_TC, RESET_AMD_APP_SSS or RESET_AMD_APP_IR"
#endif

/* triangle-count/hashtable.cu:103-184 (and the SSS / IR copies): copy the
 * app's CSR to the device, build both layouts, publish the globals. */
void make_hash_table(float load_factor, int bucket_size);
/* Releases the tables and device copies (no counterpart: the reference
 * leaves them allocated). */
void free_hash_table();
/* The handle behind the published pointers, for the C ABI's count / verify /
 * query entry points (NULL before make_hash_table). */
rs_tables* reset_amd_standalone_tables();

#endif /* RESET_AMD_STANDALONE_HPP */
