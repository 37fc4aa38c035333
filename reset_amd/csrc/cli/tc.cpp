// tc -- triangle counting over a dataset folder, the command line of
// triangle-count/TC.cu:192-347: reads csr_offsets_tri.bin, csr_cols_tri.bin,
// num_nodes_tri.bin, num_edges_tri.bin (an acyclically oriented CSR), counts
// with the normal and the hierarchical layout and prints both.  As in
// baseline/tc/main.cu:46-160, --method=original,cuco selects RESET's tables
// and/or the linear-probing (cuCollections-style) competitor.
#include "common.hpp"

int main(int argc, char** argv) {
  CliArgs a;
  if (!parse_cli(argc, argv, a, {"original", "cuco"})) return 1;
  try {
    using namespace reset_amd;
    std::vector<int> cols = read_i32_vec(a.folder + "/csr_cols_tri.bin");
    std::vector<int> offsets = read_i32_vec(a.folder + "/csr_offsets_tri.bin");
    const int num_edges = read_int_binary(a.folder + "/num_edges_tri.bin");
    const int num_nodes = read_int_binary(a.folder + "/num_nodes_tri.bin");
    std::printf("Graph loaded: %d nodes, %d edges\n", num_nodes, num_edges);
    if ((int64_t)offsets.size() != (int64_t)num_nodes + 1 || (int64_t)cols.size() != num_edges)
      throw std::runtime_error("node/edge counts do not match the CSR files");
    CSRGraph graph(offsets, cols);
    bool have = false;
    unsigned long long want = 0;
    if (a.has("original")) {
      TCBaseline tc;
      tc.build_hash_tables(num_nodes, offsets, cols, a.load_factor, a.bucket_size);
      tc.reorder_csr_by_hash_layout(graph, a.bucket_size);
      const auto n = tc.run_normal(graph, 4, 0, 0, a.bucket_size, false);
      std::printf("Normal hash - Triangles: %llu, Time: %.6f sec\n", n.first, n.second / 1000.0);
      const auto h = tc.run_hierarchical(graph, 4, 0, 0, a.bucket_size, true);
      std::printf("Hierarchical hash - Triangles: %llu, Time: %.6f sec\n", h.first,
                  h.second / 1000.0);
      if (n.first != h.first) {
        std::printf("Wrong Answer!\n");
        return 2;
      }
      have = true;
      want = n.first;
    }
    if (a.has("cuco")) {
      const auto c = run_tc_cuco(num_nodes, num_edges, nullptr, graph.get_device_offsets(),
                                 graph.get_device_elements(), offsets, cols, 0, 0, 4, a.load_factor);
      std::printf("[cuCollections] Triangles: %llu, Time: %.6f sec\n", c.first, c.second / 1000.0);
      if (have && c.first != want) {
        std::printf("Wrong Answer!\n");
        return 2;
      }
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
