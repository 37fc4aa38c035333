// solver.hip -- the analytic (load factor, bucket size) cost model of the
// reference's hyperparameter solver (hyperparameter_solver/src/solver.py),
// restated on the host in double precision.
//
// For a table of arr_size entries at load factor alpha with b-slot buckets:
//   u[j]    binomial(n = arr_size, p = alpha*b/n) probability that j keys
//           hash to a bucket, for j <= k_max = 256, the tail mass folded into
//           u[k_max] (solver.py:224-255);
//   pi[k]   stationary distribution of the overflow k a bucket passes on,
//           by fixed-point iteration of pi <- T pi (normalised) with
//           T[0][j] = sum_{i <= b-j} u[i], T[k][j] = u[k+b-j] (solver.py:258-
//           312), stopping when the largest change drops below 1e-6;
//   psi     = pi * u (convolution, first k_max+1 terms) -- bucket occupancy;
//   P_full  = sum_{m >= b} psi[m],  E_S = sum m psi[m];
//   E_L     = (b P_full + sum_{m < b} (m+1) psi[m]) / (1 - P_full);
//   T_step  = ceil(W S_slot / (E_S S_trans)) + 1;
//   cost    = E_L * T_step (1e10 if P_full >= 1 or E_S <= 0, 1e9 if
//             P_full > 0.95).
// The recommendation is the grid point alpha in {0.10, 0.15, ..., 1.00},
// b in {1..9} of least cost (c_api.py:288-338).
#include <atomic>
#include <cmath>
#include <thread>
#include <vector>

#include "common.hpp"

namespace {

constexpr int kKmax = 256;
constexpr double kConvergence = 1e-6;
constexpr int kMaxIter = 10000;

std::vector<double> binomial_probs(int64_t n, double p) {
  std::vector<double> u(kKmax + 1, 0.0);
  if (n <= 0 || p <= 0.0) {
    u[0] = 1.0;
    return u;
  }
  if (p >= 1.0) {
    if (n <= kKmax) u[(size_t)n] = 1.0;
    return u;
  }
  const double q = 1.0 - p;
  u[0] = std::pow(q, (double)n);
  const int64_t up_to = n < kKmax ? n : kKmax;
  for (int64_t j = 1; j <= up_to; ++j) u[(size_t)j] = u[(size_t)j - 1] * (double)(n - j + 1) / (double)j * (p / q);
  double s = 0.0;
  for (int64_t j = 0; j <= up_to; ++j) s += u[(size_t)j];
  const double rest = 1.0 - s > 0.0 ? 1.0 - s : 0.0;
  if (kKmax < n)
    u[kKmax] += rest;
  else
    u[(size_t)up_to] += rest;
  return u;
}

// Stationary overflow distribution; returns the iterate held when the
// change first drops below the tolerance (the one before the last product).
std::vector<double> overflow_distribution(int b, const std::vector<double>& u) {
  const int K = kKmax + 1;
  std::vector<double> prefix(K + 1, 0.0);
  for (int i = 0; i < K; ++i) prefix[i + 1] = prefix[i] + u[i];
  // T row k holds columns 0..min(k+b, k_max)
  std::vector<double> T((size_t)K * K, 0.0);
  for (int j = 0; j <= (b < kKmax ? b : kKmax); ++j) T[j] = prefix[b - j + 1];
  for (int k = 1; k <= kKmax; ++k) {
    const int mj = k + b < kKmax ? k + b : kKmax;
    for (int j = 0; j <= mj; ++j) {
      const int idx = k + b - j;
      if (idx < K) T[(size_t)k * K + j] = u[idx];
    }
  }
  // u[i] is exactly 0 beyond its last nonzero entry, so row k >= 1 of T is
  // zero left of column k + b - last; skipping those zero terms leaves every
  // sum unchanged
  int last = kKmax;
  while (last > 0 && u[last] == 0.0) --last;
  std::vector<double> pi_old(K, 0.0), pi_new(K, 0.0);
  pi_old[0] = 1.0;
  for (int it = 0; it < kMaxIter; ++it) {
    double norm = 0.0;
    for (int k = 0; k < K; ++k) {
      const int mj = k == 0 ? (b < kKmax ? b : kKmax) : (k + b < kKmax ? k + b : kKmax);
      const int j0 = k == 0 || k + b - last < 0 ? 0 : k + b - last;
      double acc = 0.0;
      for (int j = j0; j <= mj; ++j) acc += T[(size_t)k * K + j] * pi_old[j];
      pi_new[k] = acc;
      norm += acc;
    }
    if (norm > 0.0)
      for (int k = 0; k < K; ++k) pi_new[k] /= norm;
    double diff = 0.0;
    for (int k = 0; k < K; ++k) {
      const double d = std::fabs(pi_new[k] - pi_old[k]);
      diff = d > diff ? d : diff;
    }
    if (diff < kConvergence) break;
    pi_old.swap(pi_new);
  }
  return pi_old;
}

}  // namespace

extern "C" {

double rs_hyper_cost(int64_t arr_size, double alpha, int32_t b, int32_t W, int32_t S_trans,
                     int32_t S_slot) {
  if (b < 1 || arr_size < 0) return 1e10;
  const double p = alpha * (double)b / (double)arr_size;
  const std::vector<double> u = binomial_probs(arr_size, p);
  const std::vector<double> pi = overflow_distribution(b, u);
  std::vector<double> psi(kKmax + 1, 0.0);
  for (int m = 0; m <= kKmax; ++m) {
    double acc = 0.0;
    for (int k = 0; k <= m; ++k) acc += pi[k] * u[m - k];
    psi[m] = acc;
  }
  double P_full = 0.0, E_S = 0.0;
  for (int m = 0; m <= kKmax; ++m) {
    if (m >= b) P_full += psi[m];
    E_S += (double)m * psi[m];
  }
  if (P_full >= 1.0 || E_S <= 0.0) return 1e10;
  if (P_full > 0.95) return 1e9;
  double num = P_full * (double)b;
  for (int m = 0; m < b && m <= kKmax; ++m) num += (double)(m + 1) * psi[m];
  const double E_L = num / (1.0 - P_full);
  const double step = std::ceil(((double)W * (double)S_slot) / (E_S * (double)S_trans)) + 1.0;
  return E_L * step;
}

rs_status rs_hyper_solve(int64_t arr_size, int32_t W, int32_t S_trans, int32_t S_slot,
                         double* h_alpha, int32_t* h_b, double* h_cost) {
  RS_REQUIRE(h_alpha && h_b && h_cost, "rs_hyper_solve: null output");
  RS_REQUIRE(arr_size > 0 && W > 0 && S_trans > 0 && S_slot > 0, "rs_hyper_solve: bad arguments");
  // the 19 x 9 grid points are independent: spread them over host threads,
  // then take the least cost in grid order (alpha, then b)
  constexpr int kPoints = 19 * 9;
  std::vector<double> cost(kPoints);
  unsigned nt = std::thread::hardware_concurrency();
  nt = nt < 1 ? 1 : (nt > 16 ? 16 : nt);
  std::atomic<int> next{0};
  auto work = [&] {
    for (int p = next++; p < kPoints; p = next++) {
      const double alpha = std::round((0.1 + (p / 9) * 0.05) * 100.0) / 100.0;
      cost[p] = rs_hyper_cost(arr_size, alpha, p % 9 + 1, W, S_trans, S_slot);
    }
  };
  std::vector<std::thread> pool;
  for (unsigned t = 1; t < nt; ++t) pool.emplace_back(work);
  work();
  for (auto& th : pool) th.join();
  double best = cost[0], best_alpha = 0.1;
  int32_t best_b = 1;
  for (int p = 1; p < kPoints; ++p)
    if (cost[p] < best) {
      best = cost[p];
      best_alpha = std::round((0.1 + (p / 9) * 0.05) * 100.0) / 100.0;
      best_b = p % 9 + 1;
    }
  *h_alpha = best_alpha;
  *h_b = best_b;
  *h_cost = best;
  return RS_OK;
}

}  // extern "C"
