// tables.hip -- building the per-set bucketed tables (a3-a6) and the
// probe-order extraction (a7).
//
// Layout (TC.cu:18, hashtable.cu:65): one int32 array of BN*b slots; slot k of
// bucket j of table t lives at T[toff[t] + j + k*BN], so each "slot plane" k
// is a contiguous BN-int run and bucket j of every table sits in plane 0 at a
// unique offset.  toff is int64 (BN*b exceeds 2^31 at R-MAT scale 24).
#include "tables.hpp"

namespace rs {

rs_status require_layout(const rs_tables* t, int32_t layout, const int32_t** table) {
  RS_REQUIRE(t, "null tables handle");
  RS_REQUIRE(layout == RS_NORMAL || layout == RS_HIERARCHICAL, "bad layout %d", layout);
  const int32_t* p = t->table(layout);
  RS_REQUIRE(p || t->BN == 0,
             "the %s layout was not built for this handle",
             layout == RS_HIERARCHICAL ? "hierarchical" : "normal");
  *table = p;
  return RS_OK;
}

// ---------------------------------------------------------------- build (a4 + a6)
// hashtable.cu:53-101's insert: CAS(-1 -> v) walks slots k = 0..b-1 of the
// home bucket, then the next bucket mod L, bounded at L*b attempts; failed
// CAS are counted as conflicts.  Here the walk runs in LDS.  Consecutive
// tables occupy consecutive buckets of every slot plane, so a batch of
// consecutive small tables is a single run of W buckets in each of the b
// planes: a workgroup fills an LDS image of those b runs with EMPTY, inserts
// the batch's elements into it with LDS atomics, and writes the b runs out
// with 16-byte stores.  HBM sees the keys once and every table slot once,
// where an insert into HBM writes a cache line back per atomic (and needs a
// separate EMPTY fill before it).  Tables too large for a batch are filled
// and inserted in HBM, one workgroup per table.
constexpr int kBuildBlock = 256;           // large tables
constexpr int kBuildLdsBlock = 512;        // batches: 8 waves
constexpr int32_t kBuildLdsSlots = 10240;  // 40 KiB per workgroup: 4 workgroups on a CU

typedef int v4i_t __attribute__((ext_vector_type(4)));

BuildGeom build_geom(int32_t b) {
  BuildGeom g;
  g.cap = (kBuildLdsSlots / b) & ~7;  // buckets of one LDS plane
  g.lsmall = 8;                       // largest power of two <= cap / 2
  while (g.lsmall * 2 <= g.cap / 2) g.lsmall *= 2;
  g.wwin = g.cap - g.lsmall;
  return g;
}

// The walk of one key from its home bucket in table `tab` (slot k of bucket j
// at tab[j + k * stride]).  Returns false if L*b attempts found no EMPTY slot.
__device__ __forceinline__ bool insert_walk(int32_t* tab, int64_t stride, uint32_t L, int32_t b,
                                            uint32_t bucket, int32_t v, uint32_t& conflicts) {
  int32_t k = 0;
  const uint32_t bound = L * (uint32_t)b;
  uint32_t tries = 0;
  while (atomicCAS(&tab[bucket + (int64_t)k * stride], kEmpty, v) != kEmpty) {
    if (++tries >= bound) return false;
    ++conflicts;
    if (++k == b) {
      k = 0;
      bucket = (bucket + 1u) & (L - 1u);
    }
  }
  return true;
}

// Batch boundaries: a batch is a run of consecutive small tables (L <=
// lsmall) whose first buckets fall in one window of wwin buckets, so that it
// spans at most wwin + lsmall = cap buckets; a table above lsmall is a batch
// of its own, built in LDS up to cap buckets and in HBM beyond (`large`).
__global__ void batch_flags_kernel(const int32_t* __restrict__ Ls, const int64_t* __restrict__ toff,
                                   int64_t n, int32_t lsmall, int32_t cap, int64_t wwin,
                                   int32_t* __restrict__ first, int32_t* __restrict__ large) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const bool alone = Ls[t] > lsmall;
    bool start = t == 0 || alone;
    if (!start) start = Ls[t - 1] > lsmall || toff[t] / wwin != toff[t - 1] / wwin;
    first[t] = start;
    large[t] = Ls[t] > cap;
  }
}

__global__ void batch_scatter_kernel(const int32_t* __restrict__ first, const int64_t* __restrict__ fpos,
                                     const int32_t* __restrict__ large, const int64_t* __restrict__ lpos,
                                     int64_t n, int32_t* __restrict__ batch_first,
                                     int32_t* __restrict__ large_list) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    if (first[t]) batch_first[fpos[t]] = (int32_t)t;
    if (large[t]) large_list[lpos[t]] = (int32_t)t;
    if (t == 0) batch_first[fpos[n]] = (int32_t)n;
  }
}

// One workgroup per batch at a time.  The image is all EMPTY when a batch
// starts: it is filled once, and the write-out puts EMPTY back behind every
// slot it copies.  The next batch's keys (up to kBuildPre per thread) are
// loaded before the write-out, so their latency hides behind the stores.
constexpr int kBuildPre = 2;

// A workgroup barrier that waits for this thread's LDS operations only: the
// write-out's global stores go to slots no later batch reads, so they need
// not drain before the next batch starts.
__device__ __forceinline__ void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

template <int LAYOUT>
__global__ void __launch_bounds__(kBuildLdsBlock) build_lds_kernel(
    const int32_t* __restrict__ batch_first, int32_t nbatches, const int32_t* __restrict__ off,
    const int32_t* __restrict__ rows, const int32_t* __restrict__ elems, const int32_t* __restrict__ Ls,
    const int64_t* __restrict__ toff, int64_t BN, int32_t b, int32_t M, int32_t cap,
    int32_t* __restrict__ T, int32_t* status, unsigned long long* conflicts) {
  extern __shared__ __attribute__((aligned(16))) int32_t img[];  // b planes of cap slots
  uint32_t conf = 0;
  bool full = false, bad = false;
  const v4i_t empty4 = {kEmpty, kEmpty, kEmpty, kEmpty};
  const int32_t tid = (int32_t)threadIdx.x;
  for (int32_t i = 4 * tid; i < cap * b; i += 4 * kBuildLdsBlock)
    *reinterpret_cast<v4i_t*>(img + i) = empty4;
  // the batch `bi`: tables [t0, t1), elements [e0, e1), buckets from `base`;
  // `small` is false for a large table (skipped, block-uniform)
  int32_t t0 = 0, t1 = 0, e0 = 0, e1 = 0;
  int64_t base = 0;
  bool small = false;
  int32_t pu[kBuildPre], pv[kBuildPre];
  auto fetch = [&](int32_t bi) {
    t0 = batch_first[bi];
    t1 = batch_first[bi + 1];
    small = Ls[t0] <= cap;
    base = toff[t0];
    e0 = off[t0];
    e1 = off[t1];
#pragma unroll
    for (int r = 0; r < kBuildPre; ++r) {
      const int32_t e = e0 + tid + r * kBuildLdsBlock;
      pu[r] = 0;
      pv[r] = 0;
      if (small && e < e1) {
        pu[r] = rows[e];
        pv[r] = elems[e];
      }
    }
  };
  auto insert = [&](int32_t u, int32_t v) {
    if (v == kEmpty) {
      bad = true;
      return;
    }
    const uint32_t L = (uint32_t)Ls[u];
    int32_t* tab = img + (int32_t)(toff[u] - base);
    full |= !insert_walk(tab, cap, L, b, home_bucket<LAYOUT>((uint32_t)v, L, (uint32_t)M), v, conf);
  };
  int32_t bi = blockIdx.x;
  if (bi < nbatches) fetch(bi);
  lds_barrier();
  while (bi < nbatches) {
    const int32_t nbi = bi + (int32_t)gridDim.x;
    if (!small) {  // a large table: built in HBM
      if (nbi < nbatches) fetch(nbi);
      bi = nbi;
      continue;
    }
#pragma unroll
    for (int r = 0; r < kBuildPre; ++r)
      if (e0 + tid + r * kBuildLdsBlock < e1) insert(pu[r], pv[r]);
    for (int32_t e = e0 + tid + kBuildPre * kBuildLdsBlock; e < e1; e += kBuildLdsBlock)
      insert(rows[e], elems[e]);
    lds_barrier();
    const int64_t obase = base;
    const int32_t W = (int32_t)(toff[t1] - obase);  // a multiple of 8, <= cap
    if (nbi < nbatches) fetch(nbi);
    for (int32_t k = 0; k < b; ++k) {
      int32_t* dst = T + obase + (int64_t)k * BN;
      for (int32_t i = 4 * tid; i < W; i += 4 * kBuildLdsBlock) {
        v4i_t* s4 = reinterpret_cast<v4i_t*>(img + k * cap + i);
        *reinterpret_cast<v4i_t*>(dst + i) = *s4;
        *s4 = empty4;
      }
    }
    lds_barrier();
    bi = nbi;
  }
  const uint32_t sc = wave_sum(conf);
  if (__lane_id() == 0 && sc) atomicAdd(conflicts, (unsigned long long)sc);
  if (__any(full) && __lane_id() == 0) atomicOr(status, kErrTableFull);
  if (__any(bad) && __lane_id() == 0) atomicOr(status, kErrBadInput);
}

// The large tables: EMPTY into their b runs, then (next kernel) their
// elements inserted in HBM; one workgroup per table.
__global__ void __launch_bounds__(kBuildBlock) fill_large_kernel(const int32_t* __restrict__ list, int32_t nlist,
                                                                 const int32_t* __restrict__ Ls,
                                                                 const int64_t* __restrict__ toff, int64_t BN,
                                                                 int32_t b, int32_t* __restrict__ T) {
  const v4i_t empty4 = {kEmpty, kEmpty, kEmpty, kEmpty};
  for (int32_t i = blockIdx.x; i < nlist; i += gridDim.x) {
    const int32_t t = list[i];
    const int32_t L = Ls[t];
    for (int32_t k = 0; k < b; ++k) {
      int32_t* dst = T + toff[t] + (int64_t)k * BN;
      for (int32_t j = 4 * (int32_t)threadIdx.x; j < L; j += 4 * kBuildBlock)
        *reinterpret_cast<v4i_t*>(dst + j) = empty4;
    }
  }
}

template <int LAYOUT>
__global__ void __launch_bounds__(kBuildBlock) insert_large_kernel(
    const int32_t* __restrict__ list, int32_t nlist, const int32_t* __restrict__ off,
    const int32_t* __restrict__ elems, const int32_t* __restrict__ Ls, const int64_t* __restrict__ toff,
    int64_t BN, int32_t b, int32_t M, int32_t* __restrict__ T, int32_t* status,
    unsigned long long* conflicts) {
  uint32_t conf = 0;
  bool full = false, bad = false;
  for (int32_t i = blockIdx.x; i < nlist; i += gridDim.x) {
    const int32_t t = list[i];
    const uint32_t L = (uint32_t)Ls[t];
    int32_t* tab = T + toff[t];
    for (int32_t e = off[t] + (int32_t)threadIdx.x; e < off[t + 1]; e += kBuildBlock) {
      const int32_t v = elems[e];
      if (v == kEmpty) {
        bad = true;
        continue;
      }
      full |= !insert_walk(tab, BN, L, b, home_bucket<LAYOUT>((uint32_t)v, L, (uint32_t)M), v, conf);
    }
  }
  const uint32_t sc = wave_sum(conf);
  if (__lane_id() == 0 && sc) atomicAdd(conflicts, (unsigned long long)sc);
  if (__any(full) && __lane_id() == 0) atomicOr(status, kErrTableFull);
  if (__any(bad) && __lane_id() == 0) atomicOr(status, kErrBadInput);
}

// baseline/common/utils.cu:76-146: one wave per set walks slot index
// s = bucket*b + slot (bucket-major) and appends non-empty slots in order,
// compacting with a 64-bit ballot.
__global__ void __launch_bounds__(256) probe_order_kernel(const int32_t* __restrict__ off, int64_t n,
                                                          const int32_t* __restrict__ Ls,
                                                          const int64_t* __restrict__ toff,
                                                          int64_t BN, int32_t b,
                                                          const int32_t* __restrict__ T,
                                                          int32_t* __restrict__ out,
                                                          int32_t* status) {
  const int lane = __lane_id();
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
  for (int64_t t = wave; t < n; t += nwaves) {
    const int32_t start = off[t], deg = off[t + 1] - start;
    if (deg == 0) continue;
    const int32_t slots = Ls[t] * b;
    const int32_t* Tt = T + toff[t];
    int32_t written = 0;
    for (int32_t s0 = 0; s0 < slots && written < deg; s0 += kWave) {
      const int32_t s = s0 + lane;
      int32_t v = kEmpty;
      if (s < slots) v = Tt[s / b + (int64_t)(s % b) * BN];
      const bool keep = v != kEmpty;
      const uint64_t m = __ballot(keep);
      const int32_t pos = written + __popcll(m & lanemask_lt());
      if (keep && pos < deg) out[start + pos] = v;
      written += __popcll(m);
    }
    if (written != deg && lane == 0) atomicOr(status, kErrBadInput);
  }
}

// The batches of a handle (once per build): the first table of every batch
// and the list of large tables.
static rs_status plan_batches(rs_tables* t, hipStream_t st) {
  t->geom = build_geom(t->b);
  t->nbatches = 0;
  t->nlarge = 0;
  if (t->n == 0) return RS_OK;
  const int64_t n = t->n;
  DevBuf first, large, fpos, lpos, tmp;
  RS_TRY(first.alloc(sizeof(int32_t) * (size_t)n));
  RS_TRY(large.alloc(sizeof(int32_t) * (size_t)n));
  RS_TRY(fpos.alloc(sizeof(int64_t) * (size_t)(n + 1)));
  RS_TRY(lpos.alloc(sizeof(int64_t) * (size_t)(n + 1)));
  const int g = grid_for(n, 256, 16);
  batch_flags_kernel<<<g, 256, 0, st>>>(t->lengths.as<int32_t>(), t->toff.as<int64_t>(), n, t->geom.lsmall,
                                        t->geom.cap, t->geom.wwin, first.as<int32_t>(), large.as<int32_t>());
  RS_HIP(hipGetLastError());
  RS_TRY(exclusive_scan_i32_to_i64(first.as<int32_t>(), n, fpos.as<int64_t>(), tmp, st));
  RS_TRY(exclusive_scan_i32_to_i64(large.as<int32_t>(), n, lpos.as<int64_t>(), tmp, st));
  int64_t counts[2] = {0, 0};
  RS_HIP(hipMemcpyAsync(&counts[0], fpos.as<int64_t>() + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  RS_HIP(hipMemcpyAsync(&counts[1], lpos.as<int64_t>() + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  RS_HIP(hipStreamSynchronize(st));
  t->nbatches = (int32_t)counts[0];
  t->nlarge = (int32_t)counts[1];
  RS_TRY(t->batch_first.alloc(sizeof(int32_t) * (size_t)(t->nbatches + 1)));
  RS_TRY(t->large_list.alloc(sizeof(int32_t) * (size_t)(t->nlarge > 0 ? t->nlarge : 1)));
  batch_scatter_kernel<<<g, 256, 0, st>>>(first.as<int32_t>(), fpos.as<int64_t>(), large.as<int32_t>(),
                                          lpos.as<int64_t>(), n, t->batch_first.as<int32_t>(),
                                          t->large_list.as<int32_t>());
  RS_HIP(hipGetLastError());
  RS_HIP(hipStreamSynchronize(st));  // the temporaries are released on return
  return RS_OK;
}

template <int LAYOUT>
static void build_layout(rs_tables* t, int32_t* T, unsigned long long* conf, hipStream_t st) {
  const BuildGeom& g = t->geom;
  if (t->nbatches > 0) {
    int blocks = device_cu_count() * 4;
    if (blocks > t->nbatches) blocks = t->nbatches;
    const size_t lds = sizeof(int32_t) * (size_t)g.cap * (size_t)t->b;
    build_lds_kernel<LAYOUT><<<blocks, kBuildLdsBlock, lds, st>>>(
        t->batch_first.as<int32_t>(), t->nbatches, t->d_off, t->rows.as<int32_t>(), t->d_elems,
        t->lengths.as<int32_t>(), t->toff.as<int64_t>(), t->BN, t->b, t->M, g.cap, T, t->d_status(), conf);
  }
  if (t->nlarge > 0) {
    int blocks = device_cu_count() * 4;
    if (blocks > t->nlarge) blocks = t->nlarge;
    fill_large_kernel<<<blocks, kBuildBlock, 0, st>>>(t->large_list.as<int32_t>(), t->nlarge,
                                                      t->lengths.as<int32_t>(), t->toff.as<int64_t>(),
                                                      t->BN, t->b, T);
    insert_large_kernel<LAYOUT><<<blocks, kBuildBlock, 0, st>>>(
        t->large_list.as<int32_t>(), t->nlarge, t->d_off, t->d_elems, t->lengths.as<int32_t>(),
        t->toff.as<int64_t>(), t->BN, t->b, t->M, T, t->d_status(), conf);
  }
}

// a4 + a6 for the layouts in `mask` (tables allocated, batches planned).
// Conflicts are added to the counters at status+16 (normal) and status+24
// (hierarchical).
static rs_status fill_and_insert(rs_tables* t, int32_t mask, hipStream_t st) {
  auto* conf = reinterpret_cast<unsigned long long*>(t->status.as<char>() + 16);
  if (mask & RS_BUILD_NORMAL) build_layout<RS_NORMAL>(t, t->t_normal.as<int32_t>(), conf, st);
  if (mask & RS_BUILD_HIERARCHICAL) build_layout<RS_HIERARCHICAL>(t, t->t_hier.as<int32_t>(), conf + 1, st);
  RS_HIP(hipGetLastError());
  return RS_OK;
}

}  // namespace rs

using namespace rs;

extern "C" {

rs_status rs_tables_build(const int32_t* d_offsets, int64_t num_sets, const int32_t* d_elements,
                          float load_factor, int32_t bucket_size, int32_t layouts,
                          rs_stream stream, rs_tables** out) {
  RS_REQUIRE(out && d_offsets && num_sets >= 0, "rs_tables_build: null argument");
  RS_REQUIRE(num_sets < (1ll << 31) - 1, "rs_tables_build: too many sets");
  RS_REQUIRE(load_factor > 0.0f && bucket_size > 0 && bucket_size <= 64,
             "rs_tables_build: bad load_factor/bucket_size");
  RS_REQUIRE(layouts >= 1 && layouts <= 3, "rs_tables_build: bad layouts mask %d", layouts);
  *out = nullptr;
  hipStream_t st = as_stream(stream);
  rs_tables* t = new rs_tables();
  auto fail = [&](rs_status s) {
    delete t;
    return s;
  };
  (void)hipGetDevice(&t->device);
  t->n = num_sets;
  t->b = bucket_size;
  t->alpha = load_factor;
  t->layouts = layouts;
  t->d_off = d_offsets;
  t->d_elems = d_elements;
  rs_status s;
  if ((s = t->status.alloc(32)) != RS_OK) return fail(s);
  if (hipMemsetAsync(t->status.p, 0, 32, st) != hipSuccess) return fail(RS_EDEVICE);
  int32_t E32 = 0;
  if (num_sets > 0) {
    if (hipMemcpyAsync(&E32, d_offsets + num_sets, sizeof(int32_t), hipMemcpyDeviceToHost, st) !=
        hipSuccess)
      return fail(RS_EDEVICE);
    if (hipStreamSynchronize(st) != hipSuccess) return fail(RS_EDEVICE);
  }
  t->E = E32;
  if (t->E < 0 || (t->E > 0 && !d_elements)) {
    set_error("rs_tables_build: bad element count %lld", (long long)t->E);
    return fail(RS_EINVAL);
  }
  // a2: lengths + max
  if ((s = t->lengths.alloc(sizeof(int32_t) * (size_t)num_sets)) != RS_OK) return fail(s);
  if ((s = rs_calculate_lengths(d_offsets, num_sets, load_factor, bucket_size,
                                t->lengths.as<int32_t>(), &t->M, stream)) != RS_OK)
    return fail(s);
  // a3: int64 offsets, BN = toff[n]
  if ((s = t->toff.alloc(sizeof(int64_t) * (size_t)(num_sets + 1))) != RS_OK) return fail(s);
  {
    DevBuf tmp;
    if ((s = exclusive_scan_i32_to_i64(t->lengths.as<int32_t>(), num_sets, t->toff.as<int64_t>(),
                                       tmp, st)) != RS_OK)
      return fail(s);
    if (num_sets == 0 && hipMemsetAsync(t->toff.p, 0, sizeof(int64_t), st) != hipSuccess)
      return fail(RS_EDEVICE);
    if (hipMemcpyAsync(&t->BN, t->toff.as<int64_t>() + num_sets, sizeof(int64_t),
                       hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return fail(RS_EDEVICE);
  }
  // a13: row of every entry
  if ((s = t->rows.alloc(sizeof(int32_t) * (size_t)t->E)) != RS_OK) return fail(s);
  if ((s = expand_rows(d_offsets, num_sets, t->E, t->rows.as<int32_t>(), st)) != RS_OK)
    return fail(s);
  // a4 + a6: tables filled with EMPTY, then the insert
  const size_t tbytes = sizeof(int32_t) * (size_t)t->BN * (size_t)bucket_size;
  if ((layouts & RS_BUILD_NORMAL) && (s = t->t_normal.alloc(tbytes)) != RS_OK) return fail(s);
  if ((layouts & RS_BUILD_HIERARCHICAL) && (s = t->t_hier.alloc(tbytes)) != RS_OK) return fail(s);
  if ((s = plan_batches(t, st)) != RS_OK) return fail(s);
  if ((s = fill_and_insert(t, layouts, st)) != RS_OK) return fail(s);
  int32_t err = 0;
  unsigned long long conf[2] = {0, 0};
  if (hipMemcpyAsync(&err, t->status.p, sizeof(err), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(conf, t->status.as<char>() + 16, sizeof(conf), hipMemcpyDeviceToHost, st) !=
          hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return fail(RS_EDEVICE);
  t->conflicts[0] = (int64_t)conf[0];
  t->conflicts[1] = (int64_t)conf[1];
  if (err & kErrBadInput) {
    set_error("rs_tables_build: an element equals the EMPTY marker -1");
    return fail(RS_EINVAL);
  }
  if (err & kErrTableFull) {
    set_error("rs_tables_build: a table has no free slot (load_factor %.3f, b %d)",
              (double)load_factor, bucket_size);
    return fail(RS_EOVERFLOW);
  }
  *out = t;
  return RS_OK;
}

rs_status rs_tables_rebuild_async(rs_tables* t, int32_t layouts, rs_stream stream) {
  RS_REQUIRE(t, "rs_tables_rebuild_async: null handle");
  RS_REQUIRE(layouts >= 1 && layouts <= 3 && (layouts & t->layouts) == layouts,
             "rs_tables_rebuild_async: layouts mask %d names a layout that was not built (built: %d)",
             layouts, t->layouts);
  return fill_and_insert(t, layouts, as_stream(stream));
}

rs_status rs_tables_timing(rs_tables* t, int32_t enable) {
  RS_REQUIRE(t, "rs_tables_timing: null handle");
  if (enable && !t->ev0) {
    RS_HIP(hipEventCreate(&t->ev0));
    RS_HIP(hipEventCreate(&t->ev1));
  }
  t->timing = enable != 0;
  return RS_OK;
}

rs_status rs_tables_kernel_ms(rs_tables* t, float* h_ms) {
  RS_REQUIRE(t && h_ms && t->ev0, "rs_tables_kernel_ms: timing was never enabled");
  RS_HIP(hipEventSynchronize(t->ev1));
  RS_HIP(hipEventElapsedTime(h_ms, t->ev0, t->ev1));
  return RS_OK;
}

rs_status rs_tables_describe(const rs_tables* t, rs_tables_desc* o) {
  RS_REQUIRE(t && o, "rs_tables_describe: null argument");
  o->num_sets = t->n;
  o->num_elements = t->E;
  o->bucket_num = t->BN;
  o->max_length = t->M;
  o->bucket_size = t->b;
  o->load_factor = t->alpha;
  o->layouts = t->layouts;
  o->d_table_normal = t->t_normal.as<int32_t>();
  o->d_table_hierarchical = t->t_hier.as<int32_t>();
  o->d_lengths = t->lengths.as<int32_t>();
  o->d_table_offsets = t->toff.as<int64_t>();
  o->d_offsets = t->d_off;
  o->d_elements = t->d_elems;
  o->conflicts_normal = t->conflicts[0];
  o->conflicts_hierarchical = t->conflicts[1];
  return RS_OK;
}

rs_status rs_tables_status(rs_tables* t, rs_stream stream) {
  RS_REQUIRE(t, "rs_tables_status: null handle");
  hipStream_t st = as_stream(stream);
  int32_t err = 0;
  RS_HIP(hipMemcpyAsync(&err, t->status.p, sizeof(err), hipMemcpyDeviceToHost, st));
  RS_HIP(hipStreamSynchronize(st));
  if (err) {
    RS_HIP(hipMemsetAsync(t->status.p, 0, sizeof(int32_t), st));
    RS_HIP(hipStreamSynchronize(st));
    if (err & (kErrProbeBound | kErrTableFull)) {
      set_error("a probe walked a whole table without finding an EMPTY slot");
      return RS_EOVERFLOW;
    }
    set_error("kernel reported malformed input (error bits %d)", err);
    return RS_EINVAL;
  }
  return RS_OK;
}

rs_status rs_tables_probe_order(const rs_tables* t, int32_t layout, int32_t* d_out,
                                rs_stream stream) {
  const int32_t* T = nullptr;
  RS_TRY(require_layout(t, layout, &T));
  RS_REQUIRE(d_out || t->E == 0, "rs_tables_probe_order: null output");
  if (t->n == 0) return RS_OK;
  hipStream_t st = as_stream(stream);
  probe_order_kernel<<<grid_for(t->n * kWave, 256, 16), 256, 0, st>>>(
      t->d_off, t->n, t->lengths.as<int32_t>(), t->toff.as<int64_t>(), t->BN, t->b, T, d_out,
      t->d_status());
  RS_HIP(hipGetLastError());
  return RS_OK;
}

}  // extern "C"
