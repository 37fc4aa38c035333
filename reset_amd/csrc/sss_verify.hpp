// sss_verify.hpp -- batched hash verification of candidate pairs, shared by
// rs_sss_verify_async and the self-join pipeline.
#pragma once

#include "tables.hpp"

namespace rs {

// Enqueue: sizes[e] = |S_u cap S_v| for every candidate, flags[e] (if not
// NULL) by the exact Jaccard rule.  Allocates its scratch in `tmp`/`W`.
rs_status sss_verify_enqueue(const rs_tables* t, int32_t layout, const int32_t* keys,
                             const int32_t* cand_u, const int32_t* cand_v, int64_t P, int32_t p,
                             int32_t q, int32_t* flags, int32_t* sizes, DevBuf& W, DevBuf& tmp,
                             hipStream_t st);

// Grouped verification (sss_group.hip): sizes[i] = |A_{cu[i]} cap B_{cv[i]}|
// for candidates sorted by cu whose build set A_{cu[i]} fits the grouped
// LDS table at load factor alpha and bucket size b in {1, 2, 4}: at most
// group_max_build(alpha, b) elements (-1: no size fits; 2047 at 1/4, 4).
// group_partition flags those candidates: sel[i] = 1.  Synchronizes st.
int32_t group_max_build(float alpha, int32_t b);
rs_status group_verify(float alpha, int32_t b, const int32_t* a_off, const int32_t* a_el,
                       const int32_t* b_off, const int32_t* b_el, const int32_t* cu, const int32_t* cv,
                       int64_t n, int32_t* sizes, int32_t* d_err, DevBuf& tmp, hipStream_t st);
rs_status group_partition(const int32_t* a_off, const int32_t* cu, int64_t n, int32_t max_build,
                          int32_t* sel, hipStream_t st);

// The verification tables' default hyperparameters (the grouped LDS table and
// the batched pair kernel's bins).
constexpr float kVerifyAlpha = 0.25f;
constexpr int32_t kVerifyB = 4;

}  // namespace rs
