// autotune.hip -- per-size-bin choice of (load factor, bucket size, chunk,
// lanes per pair) for the batched pair kernel, by timing candidates on a
// sample (the unroll is the caller's, the same in every bin).
//
// Candidates are the settings with b in {1, 2, 4} whose LDS table
// (calculate_length(chunk) x b slots) fits one wave's 2048-slot region, in
// order of the analytic cost model's estimate (the reference's
// hyperparameter solver, restated in solver.hip), cheapest first; for a bin
// that allows it, the best few also run on 32 and 16 lanes per pair (the
// reference's probe width W).  Each bin takes the fastest measured one if it
// beats the bin's current setting by 3%, else keeps that setting.
#include <algorithm>
#include <vector>

#include "common.hpp"

namespace rs {
bool pairs_subwave_fits(const rs_pairs_opts* opts, int bin, const rs_pairs_bin* b);
rs_status pairs_launch_ordered(const rs_pairs_opts* opts, int bin, const int64_t* d_a_off,
                               const int32_t* d_a, const int64_t* d_b_off, const int32_t* d_b,
                               const int32_t* d_order, int64_t count, int32_t* d_ctl,
                               int32_t* d_sizes, hipStream_t st);
int pairs_ctl_words();
int pairs_ctl_status_index();
}  // namespace rs

using namespace rs;

namespace {
constexpr float kMargin = 0.97f;  // a candidate must be >= 3% faster than the current setting

struct Cand {
  rs_pairs_bin bin;
  double est;
};

// The prior: the reference's analytic cost model (rs_hyper_cost, the
// expected probe steps times memory steps per lookup) for a table of `chunk`
// keys, with gfx950's LDS in place of the reference's GPU parameters -- 64
// lanes, 128-byte LDS transactions, 4-byte slots.  Computed once.
const std::vector<Cand>& candidates() {
  static const std::vector<Cand> cached = [] {
    const float alphas[] = {0.0625f, 0.125f, 0.25f, 0.5f, 0.75f, 0.9f};
    const int bs[] = {1, 2, 4};
    const int chunks[] = {127, 255, 511, 1024};
    std::vector<Cand> out;
    for (float a : alphas)
      for (int b : bs)
        for (int c : chunks) {
          const int32_t L = calc_length(c, a, b);
          if (L <= 0 || (int64_t)L * b > 2048) continue;  // default LDS region
          out.push_back(Cand{rs_pairs_bin{a, b, c, 0, 64}, rs_hyper_cost(c, a, b, 64, 128, 4)});
        }
    // cheapest first; among equal costs the larger chunk (fewer rebuilds)
    std::stable_sort(out.begin(), out.end(), [](const Cand& x, const Cand& y) {
      return x.est < y.est || (x.est == y.est && x.bin.chunk > y.bin.chunk);
    });
    if (out.size() > 16) out.resize(16);
    return out;
  }();
  return cached;
}
}  // namespace

extern "C" rs_status rs_pairs_autotune(const int64_t* d_a_off, const int32_t* d_a,
                                       const int64_t* d_b_off, const int32_t* d_b,
                                       int64_t num_pairs, int32_t reps, rs_pairs_opts* opts,
                                       rs_stream stream) {
  RS_REQUIRE(d_a_off && d_b_off && opts && num_pairs > 0 && reps > 0, "rs_pairs_autotune: bad args");
  hipStream_t st = as_stream(stream);
  std::vector<int64_t> ha(num_pairs + 1), hb(num_pairs + 1);
  RS_HIP(hipMemcpyAsync(ha.data(), d_a_off, sizeof(int64_t) * ha.size(), hipMemcpyDeviceToHost, st));
  RS_HIP(hipMemcpyAsync(hb.data(), d_b_off, sizeof(int64_t) * hb.size(), hipMemcpyDeviceToHost, st));
  RS_HIP(hipStreamSynchronize(st));
  const int nb = opts->num_bins;
  std::vector<std::vector<int32_t>> ids(nb);
  for (int64_t i = 0; i < num_pairs; ++i) {
    const int64_t tot = (ha[i + 1] - ha[i]) + (hb[i + 1] - hb[i]);
    int bin = tot <= 1 ? 0 : 63 - __builtin_clzll((unsigned long long)tot);
    bin = bin >= nb ? nb - 1 : bin;
    ids[bin].push_back((int32_t)i);
  }
  DevBuf order, ctl, sizes;
  RS_TRY(order.alloc(sizeof(int32_t) * (size_t)num_pairs));
  RS_TRY(ctl.alloc(sizeof(int32_t) * (size_t)pairs_ctl_words()));
  RS_TRY(sizes.alloc(sizeof(int32_t) * (size_t)num_pairs));
  RS_HIP(hipMemsetAsync(ctl.p, 0, sizeof(int32_t) * (size_t)pairs_ctl_words(), st));
  hipEvent_t e0, e1;
  RS_HIP(hipEventCreate(&e0));
  RS_HIP(hipEventCreate(&e1));
  const auto& cands = candidates();
  rs_status result = RS_OK;
  for (int bin = 0; bin < nb && result == RS_OK; ++bin) {
    if (ids[bin].size() < 16) continue;
    const int64_t cnt = (int64_t)ids[bin].size();
    if (hipMemcpyAsync(order.p, ids[bin].data(), sizeof(int32_t) * (size_t)cnt,
                       hipMemcpyHostToDevice, st) != hipSuccess) {
      result = RS_EDEVICE;
      break;
    }
    float best = 1e30f, current_ms = 1e30f;
    rs_pairs_bin best_bin = opts->bins[bin];
    best_bin.unroll = opts->bins[0].unroll;
    const rs_pairs_bin current = best_bin;
    // the bin's current setting competes too, so tuning never picks a slower
    // one; a candidate must beat it by kMargin to replace it, so timing noise
    // does not flip the choice from run to run
    std::vector<Cand> tries;
    tries.push_back(Cand{best_bin, 0.0});
    tries.insert(tries.end(), cands.begin(), cands.end());
    // part-wave probe widths for the bin's current setting and the model's
    // three cheapest, where the bin allows them
    const size_t nfull = tries.size();
    for (size_t k = 0; k < nfull && k < 4; ++k)
      for (int lanes : {32, 16}) {
        Cand c = tries[k];
        c.bin.lanes = lanes;
        if (pairs_subwave_fits(opts, bin, &c.bin)) tries.push_back(c);
      }
    for (const Cand& c : tries) {
      // only this bin's kernel runs; the other bins keep their (valid) settings
      rs_pairs_opts trial = *opts;
      trial.bins[bin] = c.bin;
      trial.bins[bin].unroll = opts->bins[0].unroll;
      float ms_min = 1e30f;
      for (int r = 0; r <= reps && result == RS_OK; ++r) {  // r == 0 warms up
        (void)hipEventRecord(e0, st);
        result = pairs_launch_ordered(&trial, bin, d_a_off, d_a, d_b_off, d_b, order.as<int32_t>(), cnt,
                                      ctl.as<int32_t>(), sizes.as<int32_t>(), st);
        (void)hipEventRecord(e1, st);
        (void)hipEventSynchronize(e1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (r > 0) ms_min = std::min(ms_min, ms);
      }
      if (&c == &tries[0]) current_ms = ms_min;
      if (ms_min < best) {
        best = ms_min;
        best_bin = c.bin;
        best_bin.unroll = opts->bins[0].unroll;
      }
    }
    opts->bins[bin] = best < kMargin * current_ms ? best_bin : current;
  }
  int32_t err = 0;
  (void)hipMemcpy(&err, ctl.as<int32_t>() + pairs_ctl_status_index(), sizeof(err), hipMemcpyDeviceToHost);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (result == RS_OK && err) {
    set_error("rs_pairs_autotune: a candidate overflowed its table");
    result = RS_EOVERFLOW;
  }
  return result;
}
