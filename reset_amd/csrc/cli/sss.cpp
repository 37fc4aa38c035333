// sss -- set-similarity verification over a dataset folder, the command line
// of set-similarity-search/sss.cu:191-337: reads vertexs.bin, csr_cols.bin,
// csr_offsets.bin, num_nodes.bin, num_edges.bin; candidate pair e is
// (vertexs[e], csr_cols[e]); counts the pairs with Jaccard >= 0.25 (the
// threshold hard-coded at sss.cu:260) under both layouts.  As in
// baseline/sss/main.cu:36-300, --method=original,cuco,cuckoo,hopscotch,
// roaring (or all) selects RESET's tables and/or the competitor tables.
#include "common.hpp"

int main(int argc, char** argv) {
  CliArgs a;
  if (!parse_cli(argc, argv, a, {"original", "cuco", "cuckoo", "hopscotch", "roaring"})) return 1;
  try {
    using namespace reset_amd;
    std::vector<int> vertexs = read_i32_vec(a.folder + "/vertexs.bin");
    std::vector<int> cols = read_i32_vec(a.folder + "/csr_cols.bin");
    std::vector<int> offsets = read_i32_vec(a.folder + "/csr_offsets.bin");
    const int num_edges = read_int_binary(a.folder + "/num_edges.bin");
    const int num_nodes = read_int_binary(a.folder + "/num_nodes.bin");
    std::printf("num_nodes: %d, num_edges: %d\n", num_nodes, num_edges);
    if ((int64_t)offsets.size() != (int64_t)num_nodes + 1 || (int64_t)cols.size() != num_edges ||
        (int64_t)vertexs.size() != num_edges)
      throw std::runtime_error("node/edge counts do not match the CSR files");
    const float threshold = 0.25f;
    CSRGraph graph(offsets, cols);
    bool have = false;
    int want = 0;
    if (a.has("original")) {
      SSSBaseline sss;
      sss.build_hash_tables(num_nodes, offsets, cols, a.load_factor, a.bucket_size, vertexs, cols);
      sss.set_vertex_pairs(vertexs);
      sss.reorder_csr_by_hash_layout(graph, a.bucket_size);
      const auto n = sss.run_normal(graph, 4, 0, 0, a.bucket_size, threshold, false);
      std::printf("[Target] Normal Kernel execution time: %.6f seconds\n", n.second / 1000.0);
      const auto h = sss.run_hierarchical(graph, 4, 0, 0, a.bucket_size, threshold, true);
      std::printf("[Target] Hierarchical Kernel execution time: %.6f seconds\n", h.second / 1000.0);
      if (n.first != h.first) {
        std::printf("Wrong Answer!\nsum1: %d, sum2: %d\n", n.first, h.first);
        return 2;
      }
      std::printf("num of pairs whose similarity is greater than threshold: %d\n", n.first);
      have = true;
      want = n.first;
    }
    using Runner = std::pair<int, float> (*)(int, int, int const*, int const*, int const*, int const*,
                                             std::vector<int> const&, std::vector<int> const&, float, int,
                                             int, int, float, hipStream_t);
    const struct {
      const char* method;
      const char* label;
      Runner run;
    } comps[] = {{"cuco", "cuCollections", run_sss_cuco},
                 {"cuckoo", "cuckoo", run_sss_cuckoo},
                 {"hopscotch", "hopscotch", run_sss_hopscotch},
                 {"roaring", "roaring", run_sss_roaring}};
    DeviceArray<int> d_vertexs;
    for (const auto& c : comps) {
      if (!a.has(c.method)) continue;
      if (!d_vertexs.size()) d_vertexs.assign(vertexs);
      const auto r = c.run(num_edges, num_nodes, d_vertexs.data(), graph.get_device_elements(),
                           graph.get_device_elements(), graph.get_device_offsets(), offsets, cols,
                           threshold, 0, 0, 4, a.load_factor, nullptr);
      std::printf("[%s] Kernel execution time: %.6f seconds\n", c.label, r.second / 1000.0);
      std::printf("[%s] num of pairs whose similarity is greater than threshold: %d\n", c.label, r.first);
      if (have && r.first != want) {
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
