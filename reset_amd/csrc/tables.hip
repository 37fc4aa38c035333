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

// hashtable.cu:53-101, both layouts in one pass over (row, element).  One
// thread per CSR entry; CAS(-1 -> v) walks slots k = 0..b-1 of the home
// bucket, then the next bucket mod L, bounded at L*b attempts.  Conflicts
// (failed CAS) are summed per wave and added once.
template <bool NORMAL, bool HIER>
__global__ void __launch_bounds__(256) build_kernel(const int32_t* __restrict__ rows,
                                                    const int32_t* __restrict__ elems, int64_t E,
                                                    const int32_t* __restrict__ Ls,
                                                    const int64_t* __restrict__ toff, int64_t BN,
                                                    int32_t b, int32_t M, int32_t* tn, int32_t* th,
                                                    int32_t* status,
                                                    unsigned long long* conflicts) {
  uint32_t cn = 0, ch = 0;
  bool full = false, bad = false;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int32_t u = rows[e];
    const int32_t v = elems[e];
    if (v == kEmpty) {
      bad = true;
      continue;
    }
    const uint32_t L = (uint32_t)Ls[u];
    const int64_t base = toff[u];
    const uint32_t bound = L * (uint32_t)b;
#pragma unroll
    for (int which = 0; which < 2; ++which) {
      if (which == 0 && !NORMAL) continue;
      if (which == 1 && !HIER) continue;
      int32_t* T = (which == 0 ? tn : th) + base;
      uint32_t bucket = which == 0 ? hash_normal((uint32_t)v, L) : hash_hier((uint32_t)v, L, (uint32_t)M);
      int32_t k = 0;
      uint32_t tries = 0;
      while (atomicCAS(&T[bucket + (int64_t)k * BN], kEmpty, v) != kEmpty) {
        if (++tries >= bound) {
          full = true;
          break;
        }
        if (which == 0) ++cn; else ++ch;
        if (++k == b) {
          k = 0;
          bucket = (bucket + 1u) & (L - 1u);
        }
      }
    }
  }
  const uint32_t sn = wave_sum(cn), sh = wave_sum(ch);
  if (__lane_id() == 0) {
    if (sn) atomicAdd(&conflicts[0], (unsigned long long)sn);
    if (sh) atomicAdd(&conflicts[1], (unsigned long long)sh);
  }
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

// Fills n4 16-byte groups with EMPTY (a kernel rather than a memset node, so
// the sequence stays capturable).
__global__ void fill_empty_kernel(int4* __restrict__ p, int64_t n4) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = int4{kEmpty, kEmpty, kEmpty, kEmpty};
}

// a4 + a6 for the layouts in `mask` (tables allocated): EMPTY fill, then the
// insert of every (row, element).  Conflicts are added to the counters at
// status+16 (normal) and status+24 (hierarchical).
static rs_status fill_and_insert(rs_tables* t, int32_t mask, hipStream_t st) {
  // BN is a multiple of 8 (every length is a power of two >= 8), so a plane
  // holds whole 16-byte groups
  const int64_t n4 = t->BN * (int64_t)t->b / 4;
  const int fg = grid_for(n4, 256, 16);
  if (mask & RS_BUILD_NORMAL) fill_empty_kernel<<<fg, 256, 0, st>>>(t->t_normal.as<int4>(), n4);
  if (mask & RS_BUILD_HIERARCHICAL) fill_empty_kernel<<<fg, 256, 0, st>>>(t->t_hier.as<int4>(), n4);
  if (t->E > 0) {
    auto* conf = reinterpret_cast<unsigned long long*>(t->status.as<char>() + 16);
    const int grid = grid_for(t->E, 256, 16);
    if (mask == RS_BUILD_BOTH)
      build_kernel<true, true><<<grid, 256, 0, st>>>(
          t->rows.as<int32_t>(), t->d_elems, t->E, t->lengths.as<int32_t>(), t->toff.as<int64_t>(),
          t->BN, t->b, t->M, t->t_normal.as<int32_t>(), t->t_hier.as<int32_t>(), t->d_status(), conf);
    else if (mask == RS_BUILD_NORMAL)
      build_kernel<true, false><<<grid, 256, 0, st>>>(
          t->rows.as<int32_t>(), t->d_elems, t->E, t->lengths.as<int32_t>(), t->toff.as<int64_t>(),
          t->BN, t->b, t->M, t->t_normal.as<int32_t>(), nullptr, t->d_status(), conf);
    else
      build_kernel<false, true><<<grid, 256, 0, st>>>(
          t->rows.as<int32_t>(), t->d_elems, t->E, t->lengths.as<int32_t>(), t->toff.as<int64_t>(),
          t->BN, t->b, t->M, nullptr, t->t_hier.as<int32_t>(), t->d_status(), conf);
  }
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
