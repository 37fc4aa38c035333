// common.hpp -- shared device/host helpers of the RESET-MI355X library.
//
// Everything here is written for gfx950 (CDNA4): 64-lane wavefronts, 64-bit
// ballot masks, LDS-resident scratch.  No CUDA compatibility layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/reset_amd.h"

namespace rs {

constexpr int kWave = 64;
constexpr int32_t kEmpty = -1;

// Error state: the last failure message of the calling host thread.
void set_error(const char* fmt, ...);
const char* last_error();

#define RS_HIP(expr)                                                         \
  do {                                                                       \
    hipError_t _e = (expr);                                                  \
    if (_e != hipSuccess) {                                                  \
      ::rs::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                      __FILE__, __LINE__);                                   \
      return RS_EDEVICE;                                                     \
    }                                                                        \
  } while (0)

#define RS_REQUIRE(cond, ...)       \
  do {                              \
    if (!(cond)) {                  \
      ::rs::set_error(__VA_ARGS__); \
      return RS_EINVAL;             \
    }                               \
  } while (0)

#define RS_TRY(expr)                 \
  do {                               \
    rs_status _s = (expr);           \
    if (_s != RS_OK) return _s;      \
  } while (0)

inline hipStream_t as_stream(rs_stream s) { return reinterpret_cast<hipStream_t>(s); }

// Device error bits, OR-ed into a per-object status word by kernels.
constexpr int32_t kErrProbeBound = 1;   // probe walked L*b slots without EMPTY
constexpr int32_t kErrTableFull = 2;    // insert found no free slot
constexpr int32_t kErrBadInput = 4;     // element == EMPTY or out-of-range id

// ---------------------------------------------------------------- length rule
// triangle-count/hashtable.cu:20-36.  (float)d / alpha is an IEEE float32
// divide (HIP's default fp32 division is correctly rounded), truncated to int;
// then the next power of two strictly above total/b, at least 8.
__host__ __device__ inline int32_t calc_length(int32_t degree, float alpha, int32_t b) {
  float q = (float)degree / alpha;
  if (!(q < 2147483648.0f)) return -1;
  int32_t total = (int32_t)q;
  int32_t len = total / b;
  if (len == 0) return 8;
  if (len >= (1 << 30)) return -1;
  uint32_t x = (uint32_t)len;
  x |= x >> 1;
  x |= x >> 2;
  x |= x >> 4;
  x |= x >> 8;
  x |= x >> 16;
  x++;
  return x < 8u ? 8 : (int32_t)x;
}

// TC.cuh:48-52 (normal) and TC.cuh:40-46 (hierarchical).  Lengths are powers
// of two, so the hierarchical shift is log2(M) - log2(L).
__device__ __forceinline__ uint32_t hash_normal(uint32_t x, uint32_t L) { return x & (L - 1u); }
__device__ __forceinline__ uint32_t hash_hier(uint32_t x, uint32_t L, uint32_t M) {
  return (x & (M - 1u)) >> (__builtin_ctz(M) - __builtin_ctz(L));
}
template <int LAYOUT>
__device__ __forceinline__ uint32_t home_bucket(uint32_t x, uint32_t L, uint32_t M) {
  if constexpr (LAYOUT == RS_HIERARCHICAL)
    return hash_hier(x, L, M);
  else
    return hash_normal(x, L);
}

// TC.cu:11-31 over the slot plane (slot k of bucket j at Tt[j + k*BN]),
// bounded at L*b slots.  Returns 1 found, 0 absent, -1 bound exhausted.
__device__ __forceinline__ int probe(const int32_t* __restrict__ Tt, int64_t BN, uint32_t bucket,
                                     uint32_t L, int32_t b, int32_t key) {
  const uint32_t mask = L - 1u;
  for (uint32_t step = 0; step < L; ++step) {
    const int32_t* p = Tt + bucket;
    for (int32_t k = 0; k < b; ++k) {
      const int32_t v = p[(int64_t)k * BN];
      if (v == key) return 1;
      if (v == kEmpty) return 0;
    }
    bucket = (bucket + 1u) & mask;
  }
  return -1;
}

// An int32 in LDS, addressed as such: offsets into a table typed this way stay
// 32-bit (a generic pointer's arithmetic is 64-bit until the address space is
// inferred, which left a 64-bit multiply-add per bucket address).
typedef __attribute__((address_space(3))) int32_t lds_i32;

// ---------------------------------------------------------------- wave helpers
__device__ __forceinline__ int lane_id() { return __lane_id(); }

template <typename T>
__device__ __forceinline__ T wave_inclusive_scan(T v) {
  const int lane = __lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    T o = __shfl_up(v, d, kWave);
    if (lane >= d) v += o;
  }
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, kWave);
  return v;
}

// Wave-wide max steps through DPP (gfx9 row shifts and row broadcasts): lanes
// with no source lane keep `ident`.
template <int CTRL, int ROWS>
__device__ __forceinline__ int32_t dpp_from(int32_t x, int32_t ident) {
  return __builtin_amdgcn_update_dpp(ident, x, CTRL, ROWS, 0xf, false);
}

// Inclusive prefix max over the lanes of the wave (lane order).
__device__ __forceinline__ int32_t wave_prefix_max(int32_t x) {
  constexpr int32_t I = INT32_MIN;
  x = max(x, dpp_from<0x111, 0xf>(x, I));  // row_shr:1
  x = max(x, dpp_from<0x112, 0xf>(x, I));  // row_shr:2
  x = max(x, dpp_from<0x114, 0xf>(x, I));  // row_shr:4
  x = max(x, dpp_from<0x118, 0xf>(x, I));  // row_shr:8
  x = max(x, dpp_from<0x142, 0xa>(x, I));  // row_bcast:15 into rows 1 and 3
  x = max(x, dpp_from<0x143, 0xc>(x, I));  // row_bcast:31 into rows 2 and 3
  return x;
}

// Unsigned max and min over the whole wave (every lane gets it), by the same
// DPP steps; the result is read from the last lane.
template <bool MAX>
__device__ __forceinline__ uint32_t wave_reduce_u32(uint32_t x) {
  const int32_t I = MAX ? 0 : -1;
  auto op = [](uint32_t a, uint32_t b) { return MAX ? (a > b ? a : b) : (a < b ? a : b); };
  x = op(x, (uint32_t)dpp_from<0x111, 0xf>((int32_t)x, I));
  x = op(x, (uint32_t)dpp_from<0x112, 0xf>((int32_t)x, I));
  x = op(x, (uint32_t)dpp_from<0x114, 0xf>((int32_t)x, I));
  x = op(x, (uint32_t)dpp_from<0x118, 0xf>((int32_t)x, I));
  x = op(x, (uint32_t)dpp_from<0x142, 0xa>((int32_t)x, I));
  x = op(x, (uint32_t)dpp_from<0x143, 0xc>((int32_t)x, I));
  return (uint32_t)__builtin_amdgcn_readlane((int32_t)x, kWave - 1);
}

// The value of lane - 1 (lane 0 gets `ident`): DPP wave_shr:1.
__device__ __forceinline__ int32_t wave_shr1(int32_t x, int32_t ident) {
  return dpp_from<0x138, 0xf>(x, ident);
}

__device__ __forceinline__ uint64_t lanemask_lt() {
  const int lane = __lane_id();
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Grid geometry: enough waves to fill 256 CUs several times over.
int device_cu_count();
inline int grid_for(int64_t items, int block, int max_blocks_per_cu = 8) {
  int64_t g = (items + block - 1) / block;
  int64_t cap = (int64_t)device_cu_count() * max_blocks_per_cu;
  if (g > cap) g = cap;
  return g < 1 ? 1 : (int)g;
}

// Cache of device blocks for one host thread.  While a thread has a cache
// installed (BlockCacheScope), DevBuf takes blocks from it and gives them back
// to it instead of calling hipMalloc/hipFree: hipFree waits for the whole
// device, so a serving thread that allocates per request would otherwise stall
// every other stream's queued work at each release.  Blocks are sized in
// powers of two (at least 4 KiB) and reused only by the thread that installed
// the cache, whose work is ordered on one stream: a block given back while
// kernels that read it are still queued is handed out again only to work
// queued behind them on that stream.  Free blocks beyond `max_free_bytes` are
// released with hipFree.
struct BlockCache {
  explicit BlockCache(size_t max_free_bytes) : cap(max_free_bytes) {}
  BlockCache(const BlockCache&) = delete;
  BlockCache& operator=(const BlockCache&) = delete;
  ~BlockCache() {
    for (int c = 0; c < kClasses; ++c)
      for (void* q : free_blocks[c]) (void)hipFree(q);
  }
  static constexpr int kClasses = 48;
  static int size_class(size_t n) {
    int c = 12;
    while (c < kClasses - 1 && ((size_t)1 << c) < n) ++c;
    return c;
  }
  size_t cap;
  size_t held = 0;  // bytes in free_blocks
  std::vector<void*> free_blocks[kClasses];
};
inline thread_local BlockCache* tl_block_cache = nullptr;

struct BlockCacheScope {
  BlockCache* prev;
  explicit BlockCacheScope(BlockCache* c) : prev(tl_block_cache) { tl_block_cache = c; }
  ~BlockCacheScope() { tl_block_cache = prev; }
};

// Device buffer owned by host objects.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  BlockCache* owner = nullptr;  // the cache the block returns to, if any
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  rs_status alloc(size_t n) {
    release();
    if (n == 0) n = 16;
    BlockCache* c = tl_block_cache;
    if (c) {
      const int k = BlockCache::size_class(n);
      if ((size_t)1 << k >= n) {
        n = (size_t)1 << k;
        if (!c->free_blocks[k].empty()) {
          p = c->free_blocks[k].back();
          c->free_blocks[k].pop_back();
          c->held -= n;
          bytes = n;
          owner = c;
          return RS_OK;
        }
      } else {
        c = nullptr;  // beyond the largest class: a plain allocation
      }
    }
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
      p = nullptr;
      set_error("hipMalloc(%zu bytes) failed: %s", n, hipGetErrorString(e));
      return RS_ENOMEM;
    }
    bytes = n;
    owner = c;
    return RS_OK;
  }
  void release() {
    if (p && owner && owner->held + bytes <= owner->cap) {
      owner->free_blocks[BlockCache::size_class(bytes)].push_back(p);
      owner->held += bytes;
    } else if (p) {
      (void)hipFree(p);
    }
    p = nullptr;
    bytes = 0;
    owner = nullptr;
  }
  template <typename T>
  T* as() const { return reinterpret_cast<T*>(p); }
};

// Exclusive scan of n int32 or int64 values into int64 out[0..n] (out[n] =
// total), on `stream`.  Temp storage comes from `tmp` (grown as needed).
rs_status exclusive_scan_i32_to_i64(const int32_t* in, int64_t n, int64_t* out, DevBuf& tmp,
                                    hipStream_t stream);
rs_status exclusive_scan_i64(const int64_t* in, int64_t n, int64_t* out, DevBuf& tmp,
                             hipStream_t stream);
// The same with caller-provided tile sums (scan_tile_count(n) int64 of them):
// no allocation.
int64_t scan_tile_count(int64_t n);
rs_status exclusive_scan_i32_to_i64_buf(const int32_t* in, int64_t n, int64_t* out, int64_t* sums,
                                        int64_t sums_cap, hipStream_t stream);

// rows[e] = the CSR row that holds entry e.
rs_status expand_rows(const int32_t* off, int64_t n, int64_t E, int32_t* rows, hipStream_t stream);

}  // namespace rs
