// ir -- conjunctive queries over a dataset folder, the command line of
// information-retrieval/ir.cu:236-432: reads inverted_index_offsets.bin,
// inverted_index.bin, inverted_index_num.bin, query.bin, query_offsets.bin,
// query_num.bin (each query's terms ordered by document frequency, term0
// first); runs both layouts and prints the total result count.  As in
// baseline/ir/main.cu:47-200, --method=original,cuco selects RESET's tables
// and/or the linear-probing (cuCollections-style) competitor.
#include "common.hpp"

int main(int argc, char** argv) {
  CliArgs a;
  if (!parse_cli(argc, argv, a, {"original", "cuco"})) return 1;
  try {
    using namespace reset_amd;
    std::vector<int> off = read_i32_vec(a.folder + "/inverted_index_offsets.bin");
    std::vector<int> docs = read_i32_vec(a.folder + "/inverted_index.bin");
    const int num_terms = read_int_binary(a.folder + "/inverted_index_num.bin");
    std::vector<int> query = read_i32_vec(a.folder + "/query.bin");
    std::vector<int> qoff = read_i32_vec(a.folder + "/query_offsets.bin");
    const int query_num = read_int_binary(a.folder + "/query_num.bin");
    std::printf("Successfully loaded data:\ninverted_index_num: %d\nquery_num: %d\n", num_terms,
                query_num);
    if ((int64_t)off.size() != (int64_t)num_terms + 1)
      throw std::runtime_error("term count does not match the index offsets");
    for (int t : query)
      if (t < 0 || t >= num_terms) throw std::runtime_error("query term out of range");
    InvertedIndex index(off, docs);
    IRBaseline ir;
    ir.set_queries(query, qoff, query_num);
    ir.allocate_result_buffers(index);
    bool have = false;
    int want = 0;
    if (a.has("original")) {
      ir.build_hash_tables(num_terms, off, docs, a.load_factor, a.bucket_size);
      ir.pre_sort_inverted_index(index, a.bucket_size);
      const auto n = ir.run_normal(2, 0, 0, a.bucket_size, false);
      std::printf("Normal Kernel execution time: %.6f seconds\n", n.second / 1000.0);
      const auto h = ir.run_hierarchical(2, 0, 0, a.bucket_size, true);
      std::printf("Hierarchical Kernel execution time: %.6f seconds\n", h.second / 1000.0);
      if (n.first != h.first) {
        std::printf("Result mismatch between normal and hierarchical!\n");
        return 2;
      }
      std::printf("Total result count: %d\n", n.first);
      have = true;
      want = n.first;
    }
    if (a.has("cuco")) {
      const auto c = run_ir_cuco(num_terms, query_num, index.get_device_elements(),
                                 index.get_device_offsets(), ir.get_d_query(), ir.get_d_query_offsets(),
                                 ir.get_d_result(), ir.get_d_result_offsets(), ir.get_d_result_count(),
                                 ir.get_d_G_index(), 2, off, docs, 0, 0, a.load_factor);
      std::printf("[cuCollections] Kernel execution time: %.6f seconds\n", c.second / 1000.0);
      std::printf("[cuCollections] Total result count: %d\n", c.first);
      if (have && c.first != want) {
        std::printf("Result mismatch between RESET and cuCollections tables!\n");
        return 2;
      }
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
