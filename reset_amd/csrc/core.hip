// core.hip -- error state, device queries, scans, row expansion and the
// per-set length rule (triangle-count/hashtable.cu:20-51).
#include <stdarg.h>

#include <mutex>

#include "common.hpp"

namespace rs {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

const char* last_error() { return g_err; }

int device_cu_count() {
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (cached[dev] == 0) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
      cus = 256;
    cached[dev] = cus;
  }
  return cached[dev];
}

// ------------------------------------------------------------------ scans
// Three-phase scan: per-tile totals, a single-block scan of the totals, and a
// rescan of every tile with its carry.  Tiles are 256 threads x 8 items.
constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

template <typename T>
__device__ T block_exclusive_scan(T v, T* lds_waves, T* total) {
  const int lane = __lane_id();
  const int wid = threadIdx.x / kWave;
  T inc = wave_inclusive_scan(v);
  if (lane == kWave - 1) lds_waves[wid] = inc;
  __syncthreads();
  if (wid == 0) {
    T w = lane < (kScanBlock / kWave) ? lds_waves[lane] : T(0);
    T wi = wave_inclusive_scan(w);
    if (lane < (kScanBlock / kWave)) lds_waves[lane] = wi - w;
    if (lane == kScanBlock / kWave - 1) lds_waves[kScanBlock / kWave] = wi;
  }
  __syncthreads();
  T out = lds_waves[wid] + inc - v;
  *total = lds_waves[kScanBlock / kWave];
  __syncthreads();
  return out;
}

template <typename Tin, typename Tacc>
__global__ void __launch_bounds__(kScanBlock) scan_tile_sums(const Tin* __restrict__ in, int64_t n,
                                                             Tacc* __restrict__ sums) {
  __shared__ Tacc lds[kScanBlock / kWave + 1];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  Tacc s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i)
    if (base + i < n) s += (Tacc)in[base + i];
  Tacc total;
  block_exclusive_scan<Tacc>(s, lds, &total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

template <typename Tacc>
__global__ void __launch_bounds__(kScanBlock) scan_sums_single(Tacc* __restrict__ sums,
                                                               int64_t nsums,
                                                               Tacc* __restrict__ grand) {
  __shared__ Tacc lds[kScanBlock / kWave + 1];
  Tacc carry = 0;
  for (int64_t base = 0; base < nsums; base += kScanBlock) {
    const int64_t i = base + threadIdx.x;
    Tacc v = i < nsums ? sums[i] : Tacc(0);
    Tacc total;
    Tacc ex = block_exclusive_scan<Tacc>(v, lds, &total);
    if (i < nsums) sums[i] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0 && grand) *grand = carry;
}

template <typename Tin, typename Tout, bool INCLUSIVE>
__global__ void __launch_bounds__(kScanBlock) scan_tiles(const Tin* __restrict__ in, int64_t n,
                                                         const Tout* __restrict__ sums,
                                                         Tout* __restrict__ out) {
  __shared__ Tout lds[kScanBlock / kWave + 1];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  Tout v[kScanItems];
  Tout s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    v[i] = base + i < n ? (Tout)in[base + i] : Tout(0);
    s += v[i];
  }
  Tout total;
  Tout run = block_exclusive_scan<Tout>(s, lds, &total) + sums[blockIdx.x];
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (INCLUSIVE) run += v[i];
    if (base + i < n) out[base + i] = run;
    if (!INCLUSIVE) run += v[i];
  }
}

template <typename T>
__global__ void scan_zero_total(T* p) { *p = T(0); }

// With caller-provided tile sums (at least scan_tile_count(n) of them).
template <typename Tin, typename Tout, bool INCLUSIVE>
static rs_status scan_impl_buf(const Tin* in, int64_t n, Tout* out, Tout* total_out, Tout* sums,
                               int64_t sums_cap, hipStream_t stream) {
  if (n <= 0) {
    if (total_out) scan_zero_total<Tout><<<1, 1, 0, stream>>>(total_out);
    RS_HIP(hipGetLastError());
    return RS_OK;
  }
  const int64_t tiles = (n + kScanTile - 1) / kScanTile;
  RS_REQUIRE(tiles < (1ll << 31), "scan of %lld items is too large", (long long)n);
  RS_REQUIRE(sums && tiles <= sums_cap, "scan of %lld items needs %lld tile sums (have %lld)",
             (long long)n, (long long)tiles, (long long)sums_cap);
  scan_tile_sums<Tin, Tout><<<(unsigned)tiles, kScanBlock, 0, stream>>>(in, n, sums);
  scan_sums_single<Tout><<<1, kScanBlock, 0, stream>>>(sums, tiles, total_out);
  scan_tiles<Tin, Tout, INCLUSIVE><<<(unsigned)tiles, kScanBlock, 0, stream>>>(in, n, sums, out);
  RS_HIP(hipGetLastError());
  return RS_OK;
}

template <typename Tin, typename Tout, bool INCLUSIVE>
static rs_status scan_impl(const Tin* in, int64_t n, Tout* out, Tout* total_out, DevBuf& tmp,
                           hipStream_t stream) {
  const int64_t tiles = n > 0 ? (n + kScanTile - 1) / kScanTile : 0;
  if (tiles > 0 && tmp.bytes < sizeof(Tout) * (size_t)tiles) RS_TRY(tmp.alloc(sizeof(Tout) * (size_t)tiles));
  return scan_impl_buf<Tin, Tout, INCLUSIVE>(in, n, out, total_out, tmp.as<Tout>(),
                                             (int64_t)(tmp.bytes / sizeof(Tout)), stream);
}

int64_t scan_tile_count(int64_t n) { return n > 0 ? (n + kScanTile - 1) / kScanTile : 0; }

rs_status exclusive_scan_i32_to_i64_buf(const int32_t* in, int64_t n, int64_t* out, int64_t* sums,
                                        int64_t sums_cap, hipStream_t stream) {
  return scan_impl_buf<int32_t, int64_t, false>(in, n, out, out + n, sums, sums_cap, stream);
}

rs_status exclusive_scan_i32_to_i64(const int32_t* in, int64_t n, int64_t* out, DevBuf& tmp,
                                    hipStream_t stream) {
  return scan_impl<int32_t, int64_t, false>(in, n, out, out + n, tmp, stream);
}

rs_status exclusive_scan_i64(const int64_t* in, int64_t n, int64_t* out, DevBuf& tmp,
                             hipStream_t stream) {
  return scan_impl<int64_t, int64_t, false>(in, n, out, out + n, tmp, stream);
}

// ------------------------------------------------------------ row expansion
// rows[e] = #{r >= 1 : off[r] <= e}: mark the first entry of every row r >= 1
// (several empty rows may share a mark) and take an inclusive prefix sum.
__global__ void mark_row_starts(const int32_t* __restrict__ off, int64_t n, int64_t E,
                                int32_t* __restrict__ mark) {
  for (int64_t r = 1 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += (int64_t)gridDim.x * blockDim.x) {
    const int32_t s = off[r];
    if (s >= 0 && s < E) atomicAdd(&mark[s], 1);
  }
}

rs_status expand_rows(const int32_t* off, int64_t n, int64_t E, int32_t* rows, hipStream_t stream) {
  if (E <= 0) return RS_OK;
  RS_HIP(hipMemsetAsync(rows, 0, sizeof(int32_t) * (size_t)E, stream));
  if (n > 1) mark_row_starts<<<grid_for(n, 256), 256, 0, stream>>>(off, n, E, rows);
  DevBuf tmp;
  RS_TRY((scan_impl<int32_t, int32_t, true>(rows, E, rows, nullptr, tmp, stream)));
  // tmp is released at scope exit; make sure the scan has consumed it.
  RS_HIP(hipStreamSynchronize(stream));
  return RS_OK;
}

// ------------------------------------------------------------ length rule
// triangle-count/hashtable.cu:39-51 with a block-level max instead of one
// atomicMax per set.  Errors (length overflow) are reported through *err.
__global__ void lengths_kernel(const int32_t* __restrict__ off, int64_t n, float alpha, int32_t b,
                               int32_t* __restrict__ L, int32_t* __restrict__ max_len,
                               int32_t* __restrict__ err) {
  int32_t m = 0;
  bool bad = false;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t d = off[i + 1] - off[i];
    const int32_t l = d < 0 ? -1 : calc_length(d, alpha, b);
    L[i] = l;
    bad |= l < 0;
    m = l > m ? l : m;
  }
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) {
    const int32_t o = __shfl_xor(m, d, kWave);
    m = o > m ? o : m;
  }
  if (__lane_id() == 0) atomicMax(max_len, m);
  if (__any(bad) && __lane_id() == 0) atomicOr(err, kErrBadInput);
}

}  // namespace rs

using namespace rs;

extern "C" {

const char* rs_version(void) { return "reset_amd 0.1 (gfx950)"; }

const char* rs_status_string(rs_status s) {
  switch (s) {
    case RS_OK: return "ok";
    case RS_EINVAL: return "invalid argument";
    case RS_ENOMEM: return "out of device memory";
    case RS_EOVERFLOW: return "hash table overflow (probe bound reached)";
    case RS_EDEVICE: return "device error";
    case RS_EBUSY: return "busy";
    case RS_ECANCELED: return "cancelled";
    case RS_ETIMEDOUT: return "timed out in the queue";
    default: return "unknown status";
  }
}

const char* rs_last_error(void) { return last_error(); }

int32_t rs_calculate_length(int32_t degree, float load_factor, int32_t bucket_size) {
  if (degree < 0 || !(load_factor > 0.0f) || bucket_size <= 0) return -1;
  return calc_length(degree, load_factor, bucket_size);
}

rs_status rs_calculate_lengths(const int32_t* d_offsets, int64_t num_sets, float load_factor,
                               int32_t bucket_size, int32_t* d_lengths, int32_t* h_max_length,
                               rs_stream stream) {
  RS_REQUIRE(d_offsets && d_lengths && num_sets >= 0, "rs_calculate_lengths: null argument");
  RS_REQUIRE(load_factor > 0.0f && bucket_size > 0, "rs_calculate_lengths: bad alpha/b");
  hipStream_t st = as_stream(stream);
  DevBuf scratch;
  RS_TRY(scratch.alloc(2 * sizeof(int32_t)));
  RS_HIP(hipMemsetAsync(scratch.p, 0, 2 * sizeof(int32_t), st));
  if (num_sets > 0)
    lengths_kernel<<<grid_for(num_sets, 256), 256, 0, st>>>(d_offsets, num_sets, load_factor,
                                                            bucket_size, d_lengths,
                                                            scratch.as<int32_t>(),
                                                            scratch.as<int32_t>() + 1);
  RS_HIP(hipGetLastError());
  int32_t h[2];
  RS_HIP(hipMemcpyAsync(h, scratch.p, sizeof(h), hipMemcpyDeviceToHost, st));
  RS_HIP(hipStreamSynchronize(st));
  if (h_max_length) *h_max_length = h[0];
  if (h[1]) {
    set_error("rs_calculate_lengths: a set's degree is negative (offsets decrease) or its length "
              "overflows int32 (degree too large)");
    return RS_EINVAL;
  }
  return RS_OK;
}

rs_status rs_expand_rows(const int32_t* d_offsets, int64_t num_sets, int64_t num_elements,
                         int32_t* d_rows, rs_stream stream) {
  RS_REQUIRE(d_offsets && (d_rows || num_elements == 0), "rs_expand_rows: null argument");
  RS_REQUIRE(num_sets >= 0 && num_elements >= 0 && num_elements < (1ll << 31),
             "rs_expand_rows: bad sizes");
  return expand_rows(d_offsets, num_sets, num_elements, d_rows, as_stream(stream));
}

}  // extern "C"
