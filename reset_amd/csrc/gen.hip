// gen.hip -- seeded synthetic inputs on the device (not on the measured
// path).  Integer-only arithmetic so every run and every device produces the
// same bits.
#include "common.hpp"

namespace rs {

__host__ __device__ inline uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__host__ __device__ inline uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// Balanced 4-round Feistel network on 24 bits (two 12-bit halves): a keyed
// bijection of [0, 2^24).
__host__ __device__ inline uint32_t feistel24(uint32_t x, uint64_t key) {
  uint32_t l = (x >> 12) & 0xFFFu, r = x & 0xFFFu;
#pragma unroll
  for (int round = 0; round < 4; ++round) {
    const uint32_t k = (uint32_t)(key >> (16 * round)) & 0xFFFFu;
    const uint32_t f = mix32(r ^ (k << 12) ^ (0x9E37u * (uint32_t)(round + 1))) & 0xFFFu;
    const uint32_t nl = r;
    r = l ^ f;
    l = nl;
  }
  return (l << 12) | r;
}

constexpr int kGenBlock = 256;
constexpr int kGenMaxBuckets = 4096;

__device__ void block_excl_scan_buckets(int32_t* hist, int32_t* start, int nbk) {
  // nbk <= 4096 = 256 threads x 16 entries
  __shared__ int32_t wsum[kGenBlock / kWave + 1];
  const int t = threadIdx.x;
  const int per = (nbk + kGenBlock - 1) / kGenBlock;
  int32_t local = 0;
  for (int q = t * per; q < (t + 1) * per && q < nbk; ++q) local += hist[q];
  const int lane = __lane_id(), wid = t / kWave;
  const int32_t inc = wave_inclusive_scan(local);
  if (lane == kWave - 1) wsum[wid] = inc;
  __syncthreads();
  if (t == 0) {
    int32_t run = 0;
    for (int w = 0; w < kGenBlock / kWave; ++w) {
      const int32_t x = wsum[w];
      wsum[w] = run;
      run += x;
    }
  }
  __syncthreads();
  int32_t run = wsum[wid] + inc - local;
  for (int q = t * per; q < (t + 1) * per && q < nbk; ++q) {
    const int32_t h = hist[q];
    start[q] = run;
    hist[q] = run;  // becomes the scatter cursor
    run += h;
  }
  __syncthreads();
}

// out[0..n) = sorted { feistel24(first + j, key) : j < n } (distinct values,
// since feistel24 is a bijection and first + n <= 2^24).
__device__ void gen_sorted_set(int32_t* __restrict__ out, int64_t n, uint32_t first, uint64_t key,
                               int32_t* hist, int32_t* start) {
  if (n <= 0) return;
  int lg = 0;
  while ((1ll << lg) < n) ++lg;
  int nbits = lg - 3;
  nbits = nbits < 0 ? 0 : (nbits > 12 ? 12 : nbits);
  const int nbk = 1 << nbits;
  const int shift = 24 - nbits;
  for (int q = threadIdx.x; q < nbk; q += kGenBlock) hist[q] = 0;
  __syncthreads();
  for (int64_t j = threadIdx.x; j < n; j += kGenBlock) {
    const uint32_t v = feistel24(first + (uint32_t)j, key);
    atomicAdd(&hist[v >> shift], 1);
  }
  __syncthreads();
  block_excl_scan_buckets(hist, start, nbk);
  for (int64_t j = threadIdx.x; j < n; j += kGenBlock) {
    const uint32_t v = feistel24(first + (uint32_t)j, key);
    const int32_t pos = atomicAdd(&hist[v >> shift], 1);
    out[pos] = (int32_t)v;
  }
  __threadfence_block();
  __syncthreads();
  for (int q = threadIdx.x; q < nbk; q += kGenBlock) {
    const int32_t s = start[q], e = hist[q];
    for (int32_t x = s + 1; x < e; ++x) {
      const int32_t v = out[x];
      int32_t y = x - 1;
      while (y >= s && out[y] > v) {
        out[y + 1] = out[y];
        --y;
      }
      out[y + 1] = v;
    }
  }
  __threadfence_block();
  __syncthreads();
}

__global__ void __launch_bounds__(kGenBlock) gen_pairs_kernel(const int64_t* __restrict__ a_off,
                                                              const int64_t* __restrict__ b_off,
                                                              const int32_t* __restrict__ kk,
                                                              int64_t P, uint64_t seed,
                                                              int32_t* a, int32_t* b) {
  __shared__ int32_t hist[kGenMaxBuckets];
  __shared__ int32_t start[kGenMaxBuckets];
  for (int64_t i = blockIdx.x; i < P; i += gridDim.x) {
    const int64_t na = a_off[i + 1] - a_off[i], nb = b_off[i + 1] - b_off[i];
    const int64_t k = kk[i];
    const uint64_t key = splitmix64(seed ^ splitmix64((uint64_t)i));
    gen_sorted_set(a + a_off[i], na, 0u, key, hist, start);
    gen_sorted_set(b + b_off[i], nb, (uint32_t)(na - k), key, hist, start);
  }
}

// R-MAT: each of the `scale` levels picks a quadrant from one 32-bit draw
// against the cumulative thresholds a, a+b, a+b+c (of 2^32); labels are then
// scrambled by x -> ((x*m1) ^ ((x*m1) >> s)) * m2 mod 2^scale (odd m1, m2:
// a bijection), so vertex id bits carry no degree information.
__global__ void gen_rmat_kernel(int32_t scale, int64_t m, uint32_t ta, uint32_t tb, uint32_t tc,
                                uint64_t seed, uint32_t m1, uint32_t m2, int32_t* __restrict__ edges) {
  const uint32_t mask = scale >= 32 ? 0xFFFFFFFFu : ((1u << scale) - 1u);
  const int sh = scale / 2 > 0 ? scale / 2 : 1;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t u = 0, v = 0;
    uint64_t z = seed ^ ((uint64_t)i * 0xD1B54A32D192ED03ull);
    for (int lvl = 0; lvl < scale; lvl += 2) {
      z = splitmix64(z);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (lvl + h >= scale) break;
        const uint32_t r = (uint32_t)(z >> (32 * h));
        const uint32_t bu = r >= ta + tb;  // quadrants c, d: u bit set
        const uint32_t bv = (r >= ta && r < ta + tb) || r >= ta + tb + tc;
        u = (u << 1) | bu;
        v = (v << 1) | bv;
      }
    }
    u = (u * m1) & mask;
    u = ((u ^ (u >> sh)) * m2) & mask;
    v = (v * m1) & mask;
    v = ((v ^ (v >> sh)) * m2) & mask;
    edges[2 * i] = (int32_t)u;
    edges[2 * i + 1] = (int32_t)v;
  }
}

}  // namespace rs

using namespace rs;

extern "C" {

rs_status rs_gen_pair_sets_async(const int64_t* d_a_off, const int64_t* d_b_off,
                                 const int32_t* d_k, int64_t num_pairs, uint64_t seed,
                                 int32_t* d_a, int32_t* d_b, rs_stream stream) {
  RS_REQUIRE(d_a_off && d_b_off && d_k && d_a && d_b, "rs_gen_pair_sets_async: null argument");
  if (num_pairs <= 0) return RS_OK;
  hipStream_t st = as_stream(stream);
  const int grid = grid_for(num_pairs * kGenBlock, kGenBlock, 8);
  gen_pairs_kernel<<<grid, kGenBlock, 0, st>>>(d_a_off, d_b_off, d_k, num_pairs, seed, d_a, d_b);
  RS_HIP(hipGetLastError());
  return RS_OK;
}

rs_status rs_gen_rmat_async(int32_t scale, int64_t num_edges, uint32_t a, uint32_t b, uint32_t c,
                            uint64_t seed, int32_t* d_edges, rs_stream stream) {
  RS_REQUIRE(scale >= 1 && scale <= 31 && num_edges >= 0 && d_edges, "rs_gen_rmat_async: bad args");
  RS_REQUIRE((uint64_t)a + b + c <= 0xFFFFFFFFull, "rs_gen_rmat_async: a+b+c exceeds 2^32");
  if (num_edges == 0) return RS_OK;
  hipStream_t st = as_stream(stream);
  const uint32_t m1 = (uint32_t)splitmix64(seed ^ 0x1234567ull) | 1u;
  const uint32_t m2 = (uint32_t)splitmix64(seed ^ 0x7654321ull) | 1u;
  gen_rmat_kernel<<<grid_for(num_edges, 256, 16), 256, 0, st>>>(scale, num_edges, a, b, c, seed,
                                                               m1, m2, d_edges);
  RS_HIP(hipGetLastError());
  return RS_OK;
}

}  // extern "C"
