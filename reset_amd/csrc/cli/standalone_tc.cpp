// standalone_tc -- a triangle-count main in the reference's free-function
// form (triangle-count/TC.cu:192-347 reads the CSR into globals, then calls
// make_hash_table), linked against tc_hashtable.o instead of hashtable.cu.
//     standalone_tc <input_folder> [--alpha=F] [--bucket=N]
// reads csr_offsets_tri.bin / csr_cols_tri.bin / num_nodes_tri.bin /
// num_edges_tri.bin, builds the tables through make_hash_table, checks the
// published globals against the handle and prints the triangle count of
// both layouts.
#include <cstdio>
#include <cstring>
#include <string>

#define RESET_AMD_APP_TC
#include "reset_amd_standalone.hpp"
#include "reset_amd.hpp"

std::vector<int> csr_row;
std::vector<int> csr_cols;
int num_nodes = 0;
int num_edges = 0;

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <input_folder> [--alpha=F] [--bucket=N]\n", argv[0]);
    return 1;
  }
  float alpha = 0.2f;
  int bucket = 5;
  for (int i = 2; i < argc; ++i) {
    if (!std::strncmp(argv[i], "--alpha=", 8)) alpha = std::stof(argv[i] + 8);
    else if (!std::strncmp(argv[i], "--bucket=", 9)) bucket = std::stoi(argv[i] + 9);
    else {
      std::fprintf(stderr, "unknown option %s\n", argv[i]);
      return 1;
    }
  }
  try {
    const std::string dir = argv[1];
    csr_row = reset_amd::read_i32_vec(dir + "/csr_offsets_tri.bin");
    csr_cols = reset_amd::read_i32_vec(dir + "/csr_cols_tri.bin");
    num_nodes = reset_amd::read_int_binary(dir + "/num_nodes_tri.bin");
    num_edges = reset_amd::read_int_binary(dir + "/num_edges_tri.bin");
    make_hash_table(alpha, bucket);
    rs_tables* t = reset_amd_standalone_tables();
    rs_tables_desc d;
    reset_amd::check(rs_tables_describe(t, &d), "describe");
    const bool same = d.d_table_normal == d_hash_table_normal &&
                      d.d_table_hierarchical == d_hash_table_hierarchical && d.d_lengths == d_hash_length &&
                      (void*)d.d_table_offsets == (void*)d_hash_tables_offset && d.max_length == max_length &&
                      d.bucket_num == bucket_num && d.d_offsets == d_csr_row && d.d_elements == d_csr_cols;
    std::printf("Max length: %d\nTotal buckets: %d\nGlobals match the handle: %s\n", max_length, bucket_num,
                same ? "yes" : "no");
    unsigned long long n0 = 0, n1 = 0;
    reset_amd::check(rs_tc_count(t, RS_NORMAL, nullptr, (uint64_t*)&n0, nullptr), "rs_tc_count");
    reset_amd::check(rs_tc_count(t, RS_HIERARCHICAL, nullptr, (uint64_t*)&n1, nullptr), "rs_tc_count");
    std::printf("Normal triangle count: %llu\nHierarchical triangle count: %llu\n", n0, n1);
    free_hash_table();
    return same && n0 == n1 ? 0 : 2;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
}
