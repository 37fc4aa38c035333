// tables.hpp -- the per-set hash table family (host-side handle).
#pragma once

#include <mutex>

#include "common.hpp"

struct rs_tc_plan;

namespace rs {
// Geometry of the LDS table build: cap buckets per LDS plane, tables of up
// to lsmall buckets are built in LDS, batches start in windows of wwin.
struct BuildGeom {
  int32_t cap = 0, lsmall = 0, wwin = 0;
};
BuildGeom build_geom(int32_t b);
}  // namespace rs

struct rs_tables {
  int device = 0;
  int64_t n = 0, E = 0, BN = 0;
  int32_t M = 0, b = 0, layouts = 0;
  float alpha = 0.f;
  const int32_t* d_off = nullptr;
  const int32_t* d_elems = nullptr;
  rs::DevBuf lengths;    // int32[n]
  rs::DevBuf toff;       // int64[n+1]
  rs::DevBuf rows;       // int32[E]  row of each CSR entry
  rs::DevBuf t_normal;   // int32[BN*b]
  rs::DevBuf t_hier;     // int32[BN*b]
  rs::DevBuf status;     // int32 error bits, then 2 x uint64 conflict counters
  rs::BuildGeom geom;          // the build's batches: first table of each
  rs::DevBuf batch_first;      // int32[nbatches + 1]
  rs::DevBuf large_list;       // int32[nlarge], tables built in HBM
  int32_t nbatches = 0, nlarge = 0;
  int64_t conflicts[2] = {0, 0};
  rs_tc_plan* tc_plan = nullptr;  // cached by rs_tc_count
  // rs_tc_count holds it from the plan's creation to the count's end: the
  // cached plan's work-stealing counters serve one count at a time
  std::mutex tc_mu;
  // rs_tables_timing: events around the kernels of the last TC count / IR
  // query on this handle
  bool timing = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  ~rs_tables() {
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
  }

  const int32_t* table(int32_t layout) const {
    return layout == RS_HIERARCHICAL ? t_hier.as<int32_t>() : t_normal.as<int32_t>();
  }
  int32_t* d_status() const { return status.as<int32_t>(); }
};

namespace rs {
// Upper bound on the candidates a similarity lookup of these tokens produces
// against the index (sss_join.hip): prefix entries sharing a rank with one of
// the query's prefix tokens.
int64_t sss_candidate_bound(const rs_sss_index* X, const int32_t* h_tokens, int32_t n);

// Checks that `layout` was built and returns the table or sets an error.
rs_status require_layout(const rs_tables* t, int32_t layout, const int32_t** table);
}  // namespace rs
