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

}  // namespace rs
