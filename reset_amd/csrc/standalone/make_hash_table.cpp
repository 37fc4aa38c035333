// make_hash_table.cpp -- the standalone apps' make_hash_table (triangle-count/
// hashtable.cu:103-184, set-similarity-search/hashtable.cu:95-197,
// information-retrieval/hashtable.cu:160-240) over the C ABI.  Compiled once
// per app with RESET_AMD_APP_TC / _SSS / _IR; see reset_amd_standalone.hpp.
#include "reset_amd_standalone.hpp"

#include <hip/hip_runtime_api.h>

#include <climits>
#include <stdexcept>
#include <string>

#include "reset_amd.hpp"

int* d_hash_table_hierarchical = nullptr;
int* d_hash_table_normal = nullptr;
int* d_hash_length = nullptr;
long long* d_hash_tables_offset = nullptr;
int max_length = 0;
#if defined(RESET_AMD_APP_TC)
int bucket_num = 0;
int* d_csr_offsets = nullptr;
int* d_csr_cols = nullptr;
int* d_csr_row = nullptr;
#elif defined(RESET_AMD_APP_SSS)
long long bucket_num = 0;
int* d_csr_offsets = nullptr;
int* d_csr_cols = nullptr;
int* d_vertexs = nullptr;
#else
long long bucket_num = 0;
int* d_inverted_index_offsets = nullptr;
int* d_inverted_index = nullptr;
#endif
int* hash_table_normal = nullptr;
int* hash_table_hierarchical = nullptr;
std::vector<int> hash_length;
std::vector<int> hash_start;

namespace {

struct Standalone {
  rs_tables* tables = nullptr;
  std::vector<void*> owned;  // device copies of the app's CSR
};

Standalone& state() {
  static Standalone s;
  return s;
}

int* to_device(const std::vector<int>& v) {
  void* p = nullptr;
  reset_amd::check_hip(hipMalloc(&p, sizeof(int) * (v.empty() ? 1 : v.size())), "make_hash_table: hipMalloc");
  state().owned.push_back(p);
  if (!v.empty())
    reset_amd::check_hip(hipMemcpy(p, v.data(), sizeof(int) * v.size(), hipMemcpyHostToDevice),
                         "make_hash_table: hipMemcpy");
  return static_cast<int*>(p);
}

}  // namespace

void free_hash_table() {
  Standalone& s = state();
  rs_tables_free(s.tables);
  s.tables = nullptr;
  for (void* p : s.owned) (void)hipFree(p);
  s.owned.clear();
  d_hash_table_hierarchical = d_hash_table_normal = d_hash_length = nullptr;
  d_hash_tables_offset = nullptr;
  max_length = 0;
  bucket_num = 0;
}

rs_tables* reset_amd_standalone_tables() { return state().tables; }

void make_hash_table(float load_factor, int bucket_size) {
  free_hash_table();
#if defined(RESET_AMD_APP_TC)
  const std::vector<int>& off = csr_row;
  const std::vector<int>& el = csr_cols;
  const int n = num_nodes;
#elif defined(RESET_AMD_APP_SSS)
  const std::vector<int>& off = csr_offsets;
  const std::vector<int>& el = csr_cols;
  const int n = num_nodes;
#else
  const std::vector<int>& off = inverted_index_offsets;
  const std::vector<int>& el = inverted_index;
  const int n = inverted_index_num;
#endif
  if (n < 0 || off.size() != (size_t)n + 1 || (size_t)off.back() != el.size())
    throw reset_amd::Error(RS_EINVAL, "make_hash_table: the CSR globals are inconsistent");
  int* d_off = to_device(off);
  int* d_el = to_device(el);
#if defined(RESET_AMD_APP_TC)
  d_csr_row = d_off;
  d_csr_offsets = d_off;
  d_csr_cols = d_el;
#elif defined(RESET_AMD_APP_SSS)
  d_csr_offsets = d_off;
  d_csr_cols = d_el;
  d_vertexs = to_device(vertexs);
#else
  d_inverted_index_offsets = d_off;
  d_inverted_index = d_el;
#endif
  rs_tables* t = nullptr;
  reset_amd::check(rs_tables_build(d_off, n, d_el, load_factor, bucket_size, RS_BUILD_BOTH, nullptr, &t),
                   "make_hash_table");
  state().tables = t;
  rs_tables_desc d;
  reset_amd::check(rs_tables_describe(t, &d), "make_hash_table");
  if (d.bucket_num > (long long)INT_MAX && sizeof(bucket_num) == sizeof(int))
    throw reset_amd::Error(RS_EOVERFLOW, "make_hash_table: " + std::to_string(d.bucket_num) +
                                             " buckets overflow the app's int bucket_num");
  d_hash_table_normal = d.d_table_normal;
  d_hash_table_hierarchical = d.d_table_hierarchical;
  d_hash_length = d.d_lengths;
  d_hash_tables_offset = reinterpret_cast<long long*>(d.d_table_offsets);
  max_length = d.max_length;
  bucket_num = (decltype(bucket_num))d.bucket_num;
}
