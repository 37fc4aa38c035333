// pairs.hip -- intersection sizes of many set pairs (set_size/intersection.cu).
//
// Plain form (intersection.cu:128-181, one pair per launch sequence): fill a
// workspace table with -1, build the larger set into it, probe the smaller.
//
// Batched form: one persistent launch over all pairs.  Each wave owns a
// private LDS table region and takes pairs from an atomic counter in the
// order of a size-binned plan (largest bins first, so the longest pairs start
// early and the small ones fill the tail).  Per pair it builds the SMALLER
// set's table in LDS -- slot-plane layout of the reference with BN = L --
// and streams the LARGER set through it once, so HBM sees each input element
// exactly once (the algorithmic minimum 4(|A|+|B|) bytes).  A build side
// larger than the bin's chunk is split into value ranges: chunk c covers
// S[c0, c0+C) and the larger set is streamed in order, stopping at the first
// element >= S[c0+C], where the next chunk's table takes over (both sets
// ascending).  Everything stays integer; the per-pair count is reduced
// across the wave and written once.
#include "common.hpp"

namespace rs {

constexpr int kPairBlock = 256;
constexpr int kWavesPerBlock = kPairBlock / kWave;
constexpr int kLdsSlotsPerWave = 2048;  // 8 KiB per wave -> 20 waves per CU
constexpr int kPlanBlock = 256;

struct PairBinDev {
  float alpha;
  int32_t b;
  int32_t chunk;
  int32_t unroll;
};

struct PairsConst {
  PairBinDev bins[RS_PAIRS_MAX_BINS];
  int32_t num_bins;
};

__device__ __forceinline__ int pair_bin(int64_t total, int num_bins) {
  int bin = total <= 1 ? 0 : 63 - __builtin_clzll((unsigned long long)total);
  return bin >= num_bins ? num_bins - 1 : bin;
}

// ---------------------------------------------------------------- plan
// Layout of the plan's control block (int32): [0..kMaxBins) histogram,
// [kMaxBins..2kMaxBins) cursors, [2kMaxBins] work counter, [2kMaxBins+1]
// error bits.  Reset by a kernel each launch (no memset inside graphs).
constexpr int kCtlHist = 0;
constexpr int kCtlCursor = RS_PAIRS_MAX_BINS;
constexpr int kCtlCounter = 2 * RS_PAIRS_MAX_BINS;
constexpr int kCtlStatus = 2 * RS_PAIRS_MAX_BINS + 1;
constexpr int kCtlWords = 2 * RS_PAIRS_MAX_BINS + 2;

__global__ void plan_reset(int32_t* ctl) {
  if (threadIdx.x < kCtlWords - 1) ctl[threadIdx.x] = 0;  // keeps the error word
}

__global__ void plan_hist(const int64_t* __restrict__ a_off, const int64_t* __restrict__ b_off,
                          int64_t P, int num_bins, int32_t* ctl) {
  __shared__ int32_t h[RS_PAIRS_MAX_BINS];
  if (threadIdx.x < RS_PAIRS_MAX_BINS) h[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < P;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t tot = (a_off[i + 1] - a_off[i]) + (b_off[i + 1] - b_off[i]);
    atomicAdd(&h[pair_bin(tot, num_bins)], 1);
  }
  __syncthreads();
  if (threadIdx.x < num_bins && h[threadIdx.x]) atomicAdd(&ctl[kCtlHist + threadIdx.x], h[threadIdx.x]);
}

// cursors[bin] = number of pairs in bins above `bin` (largest bins first)
__global__ void plan_bins_scan(int num_bins, int32_t* ctl) {
  if (threadIdx.x == 0) {
    int32_t run = 0;
    for (int bin = num_bins - 1; bin >= 0; --bin) {
      ctl[kCtlCursor + bin] = run;
      run += ctl[kCtlHist + bin];
    }
  }
}

__global__ void plan_scatter(const int64_t* __restrict__ a_off, const int64_t* __restrict__ b_off,
                             int64_t P, int num_bins, int32_t* ctl, int32_t* __restrict__ order) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < P;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t tot = (a_off[i + 1] - a_off[i]) + (b_off[i + 1] - b_off[i]);
    const int32_t pos = atomicAdd(&ctl[kCtlCursor + pair_bin(tot, num_bins)], 1);
    order[pos] = (int32_t)i;
  }
}

// ---------------------------------------------------------------- LDS table
__device__ __forceinline__ void lds_clear(int32_t* tab, int32_t slots, int lane) {
  for (int32_t s = lane; s < slots; s += kWave) tab[s] = kEmpty;
}

// hashtable.cu:53-101 inside one wave's LDS region, BN = L.
__device__ __forceinline__ bool lds_insert(int32_t* tab, uint32_t L, int32_t b, int32_t v) {
  uint32_t bucket = v & (L - 1u);
  int32_t k = 0;
  const uint32_t bound = L * (uint32_t)b;
  for (uint32_t tries = 0; tries < bound; ++tries) {
    if (atomicCAS(&tab[bucket + (uint32_t)k * L], kEmpty, v) == kEmpty) return true;
    if (++k == b) {
      k = 0;
      bucket = (bucket + 1u) & (L - 1u);
    }
  }
  return false;
}

// TC.cu:11-31 inside LDS.  1 found, 0 absent, -1 bound exhausted.
__device__ __forceinline__ int lds_probe(const int32_t* tab, uint32_t L, int32_t b, int32_t key) {
  uint32_t bucket = key & (L - 1u);
  int32_t k = 0;
  const uint32_t bound = L * (uint32_t)b;
  for (uint32_t s = 0; s < bound; ++s) {
    const int32_t v = tab[bucket + (uint32_t)k * L];
    if (v == key) return 1;
    if (v == kEmpty) return 0;
    if (++k == b) {
      k = 0;
      bucket = (bucket + 1u) & (L - 1u);
    }
  }
  return -1;
}

// One pair inside one wave.  Returns this lane's hit count.
template <int UNROLL>
__device__ __forceinline__ uint32_t pair_intersect_wave(const int32_t* __restrict__ S, int64_t ns,
                                                        const int32_t* __restrict__ Gp, int64_t ng,
                                                        int32_t* tab, const PairBinDev bin,
                                                        bool& bad) {
  const int lane = __lane_id();
  uint32_t hits = 0;
  int64_t gpos = 0;
  for (int64_t c0 = 0; c0 < ns; c0 += bin.chunk) {
    const int64_t cl = ns - c0 < bin.chunk ? ns - c0 : bin.chunk;
    const bool last = c0 + cl >= ns;
    const int64_t hi = last ? (int64_t)1 << 40 : (int64_t)S[c0 + cl];
    const uint32_t L = (uint32_t)calc_length((int32_t)cl, bin.alpha, bin.b);
    lds_clear(tab, (int32_t)(L * (uint32_t)bin.b), lane);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    for (int64_t j = lane; j < cl; j += kWave) bad |= !lds_insert(tab, L, bin.b, S[c0 + j]);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    // stream the larger set from gpos while its elements are < hi
    while (gpos < ng) {
      int32_t x[UNROLL];
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) {
        const int64_t j = gpos + u * kWave + lane;
        x[u] = j < ng ? Gp[j] : 0x7fffffff;
      }
      int64_t stop = -1;
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) {
        const int64_t j = gpos + u * kWave + lane;
        const bool in = j < ng && (int64_t)x[u] < hi;
        if (in) {
          const int r = lds_probe(tab, L, bin.b, x[u]);
          hits += r == 1;
          bad |= r < 0;
        }
        const uint64_t beyond = __ballot(j < ng && !((int64_t)x[u] < hi));
        if (stop < 0 && beyond) stop = gpos + u * kWave + (__ffsll((unsigned long long)beyond) - 1);
      }
      if (stop >= 0) {
        gpos = stop;
        break;
      }
      gpos += (int64_t)UNROLL * kWave;
    }
    __builtin_amdgcn_wave_barrier();
  }
  return hits;
}

template <int UNROLL>
__device__ __forceinline__ uint32_t pair_dispatch(const int32_t* S, int64_t ns, const int32_t* Gp,
                                                  int64_t ng, int32_t* tab, const PairBinDev& bin,
                                                  bool& bad) {
  return pair_intersect_wave<UNROLL>(S, ns, Gp, ng, tab, bin, bad);
}

__global__ void __launch_bounds__(kPairBlock) pairs_batched_kernel(
    const int64_t* __restrict__ a_off, const int32_t* __restrict__ a,
    const int64_t* __restrict__ b_off, const int32_t* __restrict__ b, int64_t P,
    const int32_t* __restrict__ order, int32_t* ctl, int32_t* __restrict__ sizes,
    const PairsConst cfg) {
  __shared__ int32_t lds[kWavesPerBlock * kLdsSlotsPerWave];
  const int lane = __lane_id();
  const int wid = threadIdx.x / kWave;
  int32_t* tab = lds + wid * kLdsSlotsPerWave;
  bool bad = false;
  while (true) {
    int32_t idx = 0;
    if (lane == 0) idx = atomicAdd(&ctl[kCtlCounter], 1);
    idx = __shfl(idx, 0, kWave);
    if (idx >= P) break;
    const int32_t i = order[idx];
    const int64_t a0 = a_off[i], a1 = a_off[i + 1], b0 = b_off[i], b1 = b_off[i + 1];
    int64_t ns = a1 - a0, ng = b1 - b0;
    const int32_t* S = a + a0;
    const int32_t* Gp = b + b0;
    if (ns > ng) {
      const int32_t* t = S;
      S = Gp;
      Gp = t;
      const int64_t tn = ns;
      ns = ng;
      ng = tn;
    }
    uint32_t hits = 0;
    if (ns > 0) {
      const PairBinDev bin = cfg.bins[pair_bin(ns + ng, cfg.num_bins)];
      switch (bin.unroll) {
        case 1: hits = pair_dispatch<1>(S, ns, Gp, ng, tab, bin, bad); break;
        case 2: hits = pair_dispatch<2>(S, ns, Gp, ng, tab, bin, bad); break;
        case 4: hits = pair_dispatch<4>(S, ns, Gp, ng, tab, bin, bad); break;
        default: hits = pair_dispatch<8>(S, ns, Gp, ng, tab, bin, bad); break;
      }
    }
    const uint32_t total = wave_sum(hits);
    if (lane == 0) sizes[i] = (int32_t)total;
  }
  if (__any(bad) && lane == 0) atomicOr(&ctl[kCtlStatus], kErrTableFull);
}

// ---------------------------------------------------------------- plain form
// intersection.cu:128-152 without the replication: every thread inserts one
// element of the build set into the global workspace table (BN = L).
__global__ void plain_build_kernel(const int32_t* __restrict__ B, int64_t nb, int32_t* tab,
                                   uint32_t L, int32_t bsz, int32_t* size_slot, int32_t* err) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j == 0) *size_slot = 0;
  if (j >= nb) return;
  const int32_t v = B[j];
  uint32_t bucket = v & (L - 1u);
  int32_t k = 0;
  const uint32_t bound = L * (uint32_t)bsz;
  for (uint32_t tries = 0;; ++tries) {
    if (tries >= bound) {
      atomicOr(err, kErrTableFull);
      return;
    }
    if (atomicCAS(&tab[bucket + (int64_t)k * L], kEmpty, v) == kEmpty) return;
    if (++k == bsz) {
      k = 0;
      bucket = (bucket + 1u) & (L - 1u);
    }
  }
}

// intersection.cu:154-181: probe each element of the smaller set, per-wave
// count, one atomic per wave.
__global__ void plain_probe_kernel(const int32_t* __restrict__ A, int64_t na,
                                   const int32_t* __restrict__ tab, uint32_t L, int32_t bsz,
                                   int32_t* size_slot, int32_t* err) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int r = 0;
  if (j < na) {
    const int32_t key = A[j];
    r = probe(tab, (int64_t)L, key & (L - 1u), L, bsz, key);
  }
  const uint32_t s = wave_sum((uint32_t)(r == 1));
  if (__lane_id() == 0 && s) atomicAdd(size_slot, (int32_t)s);
  if (__any(r < 0) && __lane_id() == 0) atomicOr(err, kErrProbeBound);
}

}  // namespace rs

using namespace rs;

struct rs_pairs_plan {
  int64_t max_pairs = 0;
  rs_pairs_opts opts{};
  PairsConst cfg{};
  DevBuf order;   // int32[max_pairs]
  DevBuf ctl;     // int32[kCtlWords]
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool timing = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;  // around the batched kernel
  ~rs_pairs_plan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
  }
};

static rs_status validate_opts(const rs_pairs_opts* o, PairsConst* cfg) {
  RS_REQUIRE(o->num_bins >= 1 && o->num_bins <= RS_PAIRS_MAX_BINS, "pairs: bad num_bins %d",
             o->num_bins);
  RS_REQUIRE(o->waves_per_cu >= kWavesPerBlock && o->waves_per_cu <= 32,
             "pairs: waves_per_cu must be in [4, 32]");
  cfg->num_bins = o->num_bins;
  for (int i = 0; i < o->num_bins; ++i) {
    const rs_pairs_bin& bn = o->bins[i];
    RS_REQUIRE(bn.load_factor > 0.f && bn.bucket_size >= 1 && bn.bucket_size <= 16 &&
                   bn.chunk >= 1,
               "pairs: bad parameters in bin %d", i);
    RS_REQUIRE(bn.unroll == 1 || bn.unroll == 2 || bn.unroll == 4 || bn.unroll == 8,
               "pairs: unroll must be 1, 2, 4 or 8 (bin %d)", i);
    const int32_t L = calc_length(bn.chunk, bn.load_factor, bn.bucket_size);
    RS_REQUIRE(L > 0 && (int64_t)L * bn.bucket_size <= kLdsSlotsPerWave,
               "pairs: bin %d table (%d buckets x %d slots) exceeds the %d-slot LDS region", i, L,
               bn.bucket_size, kLdsSlotsPerWave);
    cfg->bins[i] = PairBinDev{bn.load_factor, bn.bucket_size, bn.chunk, bn.unroll};
  }
  return RS_OK;
}

static rs_status pairs_enqueue(rs_pairs_plan* plan, const int64_t* d_a_off, const int32_t* d_a,
                               const int64_t* d_b_off, const int32_t* d_b, int64_t P,
                               int32_t* d_sizes, hipStream_t st) {
  int32_t* ctl = plan->ctl.as<int32_t>();
  int32_t* order = plan->order.as<int32_t>();
  const int nb = plan->cfg.num_bins;
  plan_reset<<<1, 64, 0, st>>>(ctl);
  if (P > 0) {
    const int g = grid_for(P, kPlanBlock, 4);
    plan_hist<<<g, kPlanBlock, 0, st>>>(d_a_off, d_b_off, P, nb, ctl);
    plan_bins_scan<<<1, 64, 0, st>>>(nb, ctl);
    plan_scatter<<<g, kPlanBlock, 0, st>>>(d_a_off, d_b_off, P, nb, ctl, order);
    const int blocks = device_cu_count() * plan->opts.waves_per_cu / kWavesPerBlock;
    if (plan->timing) RS_HIP(hipEventRecord(plan->ev0, st));
    pairs_batched_kernel<<<blocks, kPairBlock, 0, st>>>(d_a_off, d_a, d_b_off, d_b, P, order, ctl,
                                                        d_sizes, plan->cfg);
    if (plan->timing) RS_HIP(hipEventRecord(plan->ev1, st));
  }
  RS_HIP(hipGetLastError());
  return RS_OK;
}

namespace rs {
// Runs the batched kernel over the pairs listed in d_order[0..count) only
// (used by the autotuner to time one size bin at a time).
rs_status pairs_launch_ordered(const rs_pairs_opts* opts, const int64_t* d_a_off,
                               const int32_t* d_a, const int64_t* d_b_off, const int32_t* d_b,
                               const int32_t* d_order, int64_t count, int32_t* d_ctl,
                               int32_t* d_sizes, hipStream_t st) {
  PairsConst cfg{};
  RS_TRY(validate_opts(opts, &cfg));
  plan_reset<<<1, 64, 0, st>>>(d_ctl);
  const int blocks = device_cu_count() * opts->waves_per_cu / kWavesPerBlock;
  pairs_batched_kernel<<<blocks, kPairBlock, 0, st>>>(d_a_off, d_a, d_b_off, d_b, count, d_order,
                                                      d_ctl, d_sizes, cfg);
  RS_HIP(hipGetLastError());
  return RS_OK;
}
int pairs_ctl_words() { return kCtlWords; }
int pairs_ctl_status_index() { return kCtlStatus; }
}  // namespace rs

extern "C" {

void rs_pairs_default_opts(rs_pairs_opts* o) {
  if (!o) return;
  memset(o, 0, sizeof(*o));
  o->num_bins = RS_PAIRS_MAX_BINS;
  o->waves_per_cu = 16;
  for (int i = 0; i < RS_PAIRS_MAX_BINS; ++i) {
    // alpha 0.25, b 2: L = nextpow2 > 2*chunk, so a 511-element chunk fills
    // the 2048-slot region at load 1/4.
    o->bins[i] = rs_pairs_bin{0.25f, 2, 511, 8};
  }
}

rs_status rs_pairs_plan_create(int64_t max_pairs, const rs_pairs_opts* opts,
                               rs_pairs_plan** out) {
  RS_REQUIRE(out && max_pairs >= 0 && max_pairs < (1ll << 31), "rs_pairs_plan_create: bad args");
  *out = nullptr;
  auto* p = new rs_pairs_plan();
  p->max_pairs = max_pairs;
  if (opts)
    p->opts = *opts;
  else
    rs_pairs_default_opts(&p->opts);
  rs_status s = validate_opts(&p->opts, &p->cfg);
  if (s == RS_OK) s = p->order.alloc(sizeof(int32_t) * (size_t)(max_pairs > 0 ? max_pairs : 1));
  if (s == RS_OK) s = p->ctl.alloc(sizeof(int32_t) * kCtlWords);
  if (s == RS_OK && hipMemset(p->ctl.p, 0, sizeof(int32_t) * kCtlWords) != hipSuccess) {
    set_error("rs_pairs_plan_create: hipMemset failed");
    s = RS_EDEVICE;
  }
  if (s != RS_OK) {
    delete p;
    return s;
  }
  *out = p;
  return RS_OK;
}

void rs_pairs_plan_free(rs_pairs_plan* p) { delete p; }

rs_status rs_pairs_intersect_batched_async(rs_pairs_plan* plan, const int64_t* d_a_off,
                                           const int32_t* d_a, const int64_t* d_b_off,
                                           const int32_t* d_b, int64_t num_pairs, int32_t* d_sizes,
                                           rs_stream stream) {
  RS_REQUIRE(plan && d_a_off && d_b_off && (d_sizes || num_pairs == 0),
             "rs_pairs_intersect_batched_async: null argument");
  RS_REQUIRE(num_pairs >= 0 && num_pairs <= plan->max_pairs,
             "rs_pairs_intersect_batched_async: %lld pairs exceed the plan's %lld",
             (long long)num_pairs, (long long)plan->max_pairs);
  return pairs_enqueue(plan, d_a_off, d_a, d_b_off, d_b, num_pairs, d_sizes, as_stream(stream));
}

rs_status rs_pairs_graph_capture(rs_pairs_plan* plan, const int64_t* d_a_off, const int32_t* d_a,
                                 const int64_t* d_b_off, const int32_t* d_b, int64_t num_pairs,
                                 int32_t* d_sizes, rs_stream stream) {
  RS_REQUIRE(plan && d_a_off && d_b_off && d_sizes, "rs_pairs_graph_capture: null argument");
  RS_REQUIRE(num_pairs >= 0 && num_pairs <= plan->max_pairs, "rs_pairs_graph_capture: too many pairs");
  RS_REQUIRE(stream != nullptr, "rs_pairs_graph_capture: needs a non-default stream");
  hipStream_t st = as_stream(stream);
  if (plan->exec) {
    (void)hipGraphExecDestroy(plan->exec);
    plan->exec = nullptr;
  }
  if (plan->graph) {
    (void)hipGraphDestroy(plan->graph);
    plan->graph = nullptr;
  }
  (void)device_cu_count();  // attribute query outside the capture
  RS_REQUIRE(!plan->timing, "rs_pairs_graph_capture: disable kernel timing before capturing");
  RS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  rs_status s = pairs_enqueue(plan, d_a_off, d_a, d_b_off, d_b, num_pairs, d_sizes, st);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(st, &g);
  if (s != RS_OK) {
    if (g) (void)hipGraphDestroy(g);
    return s;
  }
  if (e != hipSuccess) {
    set_error("hipStreamEndCapture failed: %s", hipGetErrorString(e));
    return RS_EDEVICE;
  }
  plan->graph = g;
  RS_HIP(hipGraphInstantiate(&plan->exec, g, nullptr, nullptr, 0));
  return RS_OK;
}

rs_status rs_pairs_graph_launch(rs_pairs_plan* plan, rs_stream stream) {
  RS_REQUIRE(plan && plan->exec, "rs_pairs_graph_launch: no captured graph");
  RS_HIP(hipGraphLaunch(plan->exec, as_stream(stream)));
  return RS_OK;
}

rs_status rs_pairs_plan_timing(rs_pairs_plan* p, int32_t enable) {
  RS_REQUIRE(p, "rs_pairs_plan_timing: null plan");
  if (enable && !p->ev0) {
    RS_HIP(hipEventCreate(&p->ev0));
    RS_HIP(hipEventCreate(&p->ev1));
  }
  p->timing = enable != 0;
  return RS_OK;
}

rs_status rs_pairs_plan_kernel_ms(rs_pairs_plan* p, float* h_ms) {
  RS_REQUIRE(p && h_ms && p->ev0, "rs_pairs_plan_kernel_ms: timing was never enabled");
  RS_HIP(hipEventSynchronize(p->ev1));
  RS_HIP(hipEventElapsedTime(h_ms, p->ev0, p->ev1));
  return RS_OK;
}

rs_status rs_pairs_plan_status(rs_pairs_plan* p, rs_stream stream) {
  RS_REQUIRE(p, "rs_pairs_plan_status: null plan");
  hipStream_t st = as_stream(stream);
  int32_t err = 0;
  RS_HIP(hipMemcpyAsync(&err, p->ctl.as<int32_t>() + kCtlStatus, sizeof(err),
                        hipMemcpyDeviceToHost, st));
  RS_HIP(hipStreamSynchronize(st));
  if (err) {
    RS_HIP(hipMemsetAsync(p->ctl.as<int32_t>() + kCtlStatus, 0, sizeof(int32_t), st));
    RS_HIP(hipStreamSynchronize(st));
    set_error("pairs: an LDS table had no free slot or a probe hit its bound");
    return RS_EOVERFLOW;
  }
  return RS_OK;
}

rs_status rs_pairs_intersect_plain_async(const int64_t* d_a_off, const int32_t* d_a,
                                         const int64_t* d_b_off, const int32_t* d_b,
                                         const int64_t* h_a_off, const int64_t* h_b_off,
                                         int64_t num_pairs, float load_factor,
                                         int32_t bucket_size, int32_t* d_workspace,
                                         int32_t* d_sizes, rs_stream stream) {
  RS_REQUIRE(d_a_off && d_b_off && h_a_off && h_b_off && d_workspace && (d_sizes || !num_pairs),
             "rs_pairs_intersect_plain_async: null argument");
  RS_REQUIRE(load_factor > 0.f && bucket_size > 0, "plain pairs: bad alpha/b");
  (void)d_a_off;
  (void)d_b_off;
  hipStream_t st = as_stream(stream);
  // error word lives after the sizes in a tiny side buffer
  static thread_local DevBuf err_buf;
  if (!err_buf.p) RS_TRY(err_buf.alloc(16));
  int32_t* err = err_buf.as<int32_t>();
  for (int64_t i = 0; i < num_pairs; ++i) {
    const int64_t a0 = h_a_off[i], na = h_a_off[i + 1] - a0;
    const int64_t b0 = h_b_off[i], nb = h_b_off[i + 1] - b0;
    const int32_t* A = d_a + a0;
    const int32_t* B = d_b + b0;
    int64_t nA = na, nB = nb;
    if (nA > nB) {  // intersection.cu:198-202: probe with the smaller
      const int32_t* t = A;
      A = B;
      B = t;
      const int64_t tn = nA;
      nA = nB;
      nB = tn;
    }
    const int32_t L = calc_length((int32_t)nB, load_factor, bucket_size);
    RS_REQUIRE(L > 0, "plain pairs: set too large");
    RS_HIP(hipMemsetAsync(d_workspace, 0xFF, sizeof(int32_t) * (size_t)L * bucket_size, st));
    const unsigned gb = (unsigned)((nB + 255) / 256);
    plain_build_kernel<<<gb ? gb : 1, 256, 0, st>>>(B, nB, d_workspace, (uint32_t)L, bucket_size,
                                                    d_sizes + i, err);
    if (nA > 0)
      plain_probe_kernel<<<(unsigned)((nA + 255) / 256), 256, 0, st>>>(
          A, nA, d_workspace, (uint32_t)L, bucket_size, d_sizes + i, err);
  }
  RS_HIP(hipGetLastError());
  return RS_OK;
}

}  // extern "C"
