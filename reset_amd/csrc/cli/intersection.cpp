// intersection -- one pair of random sets, the command line of
// set_size/intersection.cu:183-303: `intersection <set A size> <set B size>
// [seed]`.  Sets are distinct values below INT_MAX from a seeded generator
// (the reference's are unseeded and may repeat values).  Times the plain
// per-pair form (workspace table of the larger set, filled, built, probed)
// and the batched LDS form, and prints the common count.
#include <algorithm>
#include <climits>
#include <cstdlib>
#include <random>
#include <unordered_set>

#include "../../../include/reset_amd.hpp"

static std::vector<int> random_set(int size, std::mt19937_64& rng) {
  std::uniform_int_distribution<int> d(0, INT_MAX - 1);
  std::unordered_set<int> seen;
  std::vector<int> v;
  v.reserve(size);
  while ((int)v.size() < size) {
    const int x = d(rng);
    if (seen.insert(x).second) v.push_back(x);
  }
  std::sort(v.begin(), v.end());
  return v;
}

int main(int argc, char** argv) {
  if (argc != 3 && argc != 4) {
    std::fprintf(stderr, "Usage: %s <set A size> <set B size> [seed]\n", argv[0]);
    return 1;
  }
  const int a_size = std::atoi(argv[1]), b_size = std::atoi(argv[2]);
  const uint64_t seed = argc == 4 ? std::strtoull(argv[3], nullptr, 10) : 20261203ull;
  if (a_size < 0 || b_size < 0) {
    std::fprintf(stderr, "set sizes must be >= 0\n");
    return 1;
  }
  try {
    using namespace reset_amd;
    std::mt19937_64 rng(seed);
    std::vector<int> a = random_set(a_size, rng), b = random_set(b_size, rng);
    // plain form: one pair, table of the larger set in a workspace
    const int big = std::max(a_size, b_size);
    const float lf = 0.2f;
    const int bs = 5;
    std::vector<int64_t> a_off = {0, (int64_t)a.size()}, b_off = {0, (int64_t)b.size()};
    DeviceArray<int64_t> da_off(a_off), db_off(b_off);
    DeviceArray<int> da(a), db(b), ws((size_t)bs * rs_calculate_length(big, lf, bs)),
        sizes(1);
    const float plain_ms = time_ms([&] {
      check(rs_pairs_intersect_plain(da_off.data(), da.data(), db_off.data(), db.data(),
                                     a_off.data(), b_off.data(), 1, lf, bs, ws.data(), sizes.data(),
                                     nullptr),
            "rs_pairs_intersect_plain");
    });
    const int plain = sizes.to_host()[0];
    std::printf("Normal Intersection Time: %f ms\n", plain_ms);
    float batched_ms = 0.f;
    const int batched = intersect_pairs(a_off, a, b_off, b, &batched_ms)[0];
    std::printf("Optimized Intersection Time: %f ms\n", batched_ms);
    if (plain != batched) {
      std::fprintf(stderr, "Mismatch in intersection results!\n");
      return 2;
    }
    std::printf("Intersection Result Count: %d\n", plain);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
