/*
 * reset_amd.hpp -- C++ mirror of the reference's host interface for the hash
 * set-intersection path, built on the C ABI of reset_amd.h.
 *
 * Same class and function names, argument meaning and result types as the
 * reference's object form (baseline/common/hash_table.cuh:12-95,
 * baseline/common/graph_data.cuh:13-97, baseline/tc/tc.cuh:29-37,
 * baseline/sss/sss.cuh:26-35, baseline/ir/ir.cuh:39-46) and its `.bin` file
 * readers (TC.cu:62-89, preprocess/util.py:12-20).  Differences a caller sees:
 *  - errors throw reset_amd::Error (a std::runtime_error carrying the
 *    rs_status) instead of printing or exiting; file readers throw
 *    std::runtime_error as baseline/common/utils.cu:11-13 does;
 *  - CHUNK_SIZE / grid_size / block_size are accepted and ignored: the
 *    kernels schedule themselves (tiled probe space, LDS-staged tables);
 *  - the float Jaccard threshold is read as the fraction p/q (q <= 4096)
 *    whose float32 value it is, and applied exactly as q*r >= p*(|u|+|v|-r).
 * Header-only; link with -lreset_amd (and the HIP runtime).
 */
#ifndef RESET_AMD_HPP
#define RESET_AMD_HPP

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "reset_amd.h"

namespace reset_amd {

struct Error : std::runtime_error {
  rs_status status;
  Error(rs_status s, const std::string& what) : std::runtime_error(what), status(s) {}
};

inline void check(rs_status s, const char* where) {
  if (s != RS_OK)
    throw Error(s, std::string(where) + ": " + rs_status_string(s) + ": " + rs_last_error());
}

inline void check_hip(hipError_t e, const char* where) {
  if (e != hipSuccess) throw Error(RS_EDEVICE, std::string(where) + ": " + hipGetErrorString(e));
}

// ------------------------------------------------------------------ files
// `.bin` vector: little-endian uint64 count, then count int32 values
// (preprocess/util.py:12-20, read by TC.cu:62-76).
inline std::vector<int> read_i32_vec(const std::string& filename) {
  std::ifstream in(filename, std::ios::binary);
  if (!in) throw std::runtime_error("Error opening file: " + filename);
  uint64_t size = 0;
  in.read(reinterpret_cast<char*>(&size), sizeof(size));
  if (!in) throw std::runtime_error("Truncated file: " + filename);
  std::vector<int> vec(size);
  in.read(reinterpret_cast<char*>(vec.data()), (std::streamsize)(size * sizeof(int)));
  if (!in) throw std::runtime_error("Truncated file: " + filename);
  return vec;
}
inline std::vector<int> read_vector_binary(const std::string& filename) {
  return read_i32_vec(filename);
}
// Scalar `.bin`: one little-endian int32 (TC.cu:78-89).
inline int read_int_binary(const std::string& filename) {
  std::ifstream in(filename, std::ios::binary);
  if (!in) throw std::runtime_error("Error opening file: " + filename);
  int value = 0;
  in.read(reinterpret_cast<char*>(&value), sizeof(value));
  if (!in) throw std::runtime_error("Truncated file: " + filename);
  return value;
}
// The 64-bit dataset files of the multi-GPU drivers (multi-GPU/sss-multi-GPU/
// main.cu:25-57): a little-endian uint64 count followed by int64 values, and a
// scalar file holding one int64.  read_vec_64 / read_num_64 narrow to int as
// the reference does, but a value outside the int32 range throws instead of
// wrapping.
inline std::vector<int64_t> read_i64_vec(const std::string& filename) {
  std::ifstream in(filename, std::ios::binary);
  if (!in) throw std::runtime_error("Error opening file: " + filename);
  uint64_t size = 0;
  in.read(reinterpret_cast<char*>(&size), sizeof(size));
  if (!in) throw std::runtime_error("Truncated file: " + filename);
  std::vector<int64_t> vec(size);
  in.read(reinterpret_cast<char*>(vec.data()), (std::streamsize)(size * sizeof(int64_t)));
  if (!in) throw std::runtime_error("Truncated file: " + filename);
  return vec;
}

inline int narrow_i32(int64_t x, const std::string& filename) {
  if (x < INT32_MIN || x > INT32_MAX)
    throw std::out_of_range("Value " + std::to_string(x) + " of " + filename + " does not fit int32");
  return (int)x;
}

inline std::vector<int> read_vec_64(const std::string& filename) {
  const std::vector<int64_t> wide = read_i64_vec(filename);
  std::vector<int> vec(wide.size());
  for (size_t i = 0; i < wide.size(); ++i) vec[i] = narrow_i32(wide[i], filename);
  return vec;
}

inline int read_num_64(const std::string& filename) {
  std::ifstream in(filename, std::ios::binary);
  if (!in) throw std::runtime_error("Error opening file: " + filename);
  int64_t value = 0;
  in.read(reinterpret_cast<char*>(&value), sizeof(value));
  if (!in) throw std::runtime_error("Truncated file: " + filename);
  return narrow_i32(value, filename);
}

inline void write_vec_64(const std::string& filename, const std::vector<int64_t>& v) {
  std::ofstream out(filename, std::ios::binary);
  if (!out) throw std::runtime_error("Error opening file: " + filename);
  const uint64_t size = v.size();
  out.write(reinterpret_cast<const char*>(&size), sizeof(size));
  out.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(size * sizeof(int64_t)));
  if (!out) throw std::runtime_error("Write failed: " + filename);
}

inline void write_num_64(const std::string& filename, int64_t value) {
  std::ofstream out(filename, std::ios::binary);
  if (!out) throw std::runtime_error("Error opening file: " + filename);
  out.write(reinterpret_cast<const char*>(&value), sizeof(value));
  if (!out) throw std::runtime_error("Write failed: " + filename);
}

inline void write_i32_vec(const std::string& filename, const std::vector<int>& v) {
  std::ofstream out(filename, std::ios::binary);
  if (!out) throw std::runtime_error("Error opening file: " + filename);
  const uint64_t size = v.size();
  out.write(reinterpret_cast<const char*>(&size), sizeof(size));
  out.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(size * sizeof(int)));
}
inline void write_int_binary(const std::string& filename, int value) {
  std::ofstream out(filename, std::ios::binary);
  if (!out) throw std::runtime_error("Error opening file: " + filename);
  out.write(reinterpret_cast<const char*>(&value), sizeof(value));
}

// ------------------------------------------------------------------ device memory
template <typename T>
class DeviceArray {
 public:
  DeviceArray() = default;
  explicit DeviceArray(size_t n) { resize(n); }
  explicit DeviceArray(const std::vector<T>& h) { assign(h); }
  ~DeviceArray() { reset(); }
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
  DeviceArray(DeviceArray&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr, o.n_ = 0; }
  DeviceArray& operator=(DeviceArray&& o) noexcept {
    if (this != &o) {
      reset();
      p_ = o.p_, n_ = o.n_;
      o.p_ = nullptr, o.n_ = 0;
    }
    return *this;
  }
  void resize(size_t n) {
    reset();
    check_hip(hipMalloc(reinterpret_cast<void**>(&p_), (n ? n : 1) * sizeof(T)), "hipMalloc");
    n_ = n;
  }
  void assign(const std::vector<T>& h) {
    resize(h.size());
    if (!h.empty())
      check_hip(hipMemcpy(p_, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
  }
  std::vector<T> to_host() const {
    std::vector<T> h(n_);
    if (n_) check_hip(hipMemcpy(h.data(), p_, n_ * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
    return h;
  }
  void reset() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* data() const { return p_; }
  size_t size() const { return n_; }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
};

// ------------------------------------------------------------------ graph data
// baseline/common/graph_data.cuh:13-97: host vectors plus device copies.
class GraphData {
 public:
  virtual ~GraphData() = default;
  virtual int get_num_nodes() const = 0;
  virtual int get_num_elements() const = 0;
  virtual const std::vector<int>& get_host_offsets() const = 0;
  virtual const std::vector<int>& get_host_elements() const = 0;
  virtual int* get_device_offsets() const = 0;
  virtual int* get_device_elements() const = 0;
};

class CSRGraph : public GraphData {
 public:
  CSRGraph(const std::string& offsets_path, const std::string& columns_path)
      : CSRGraph(read_i32_vec(offsets_path), read_i32_vec(columns_path)) {}
  CSRGraph(std::vector<int> offsets, std::vector<int> cols)
      : csr_offsets_(std::move(offsets)), csr_cols_(std::move(cols)) {
    if (csr_offsets_.empty() || (size_t)csr_offsets_.back() != csr_cols_.size())
      throw std::invalid_argument("CSRGraph: offsets do not match the column count");
    d_offsets_.assign(csr_offsets_);
    d_cols_.assign(csr_cols_);
  }
  int get_num_nodes() const override { return (int)csr_offsets_.size() - 1; }
  int get_num_elements() const override { return (int)csr_cols_.size(); }
  const std::vector<int>& get_host_offsets() const override { return csr_offsets_; }
  const std::vector<int>& get_host_elements() const override { return csr_cols_; }
  int* get_device_offsets() const override { return d_offsets_.data(); }
  int* get_device_elements() const override { return d_cols_.data(); }

 private:
  std::vector<int> csr_offsets_, csr_cols_;
  DeviceArray<int> d_offsets_, d_cols_;
};

class InvertedIndex : public GraphData {
 public:
  InvertedIndex(const std::string& index_offsets_path, const std::string& index_data_path)
      : InvertedIndex(read_i32_vec(index_offsets_path), read_i32_vec(index_data_path)) {}
  InvertedIndex(std::vector<int> offsets, std::vector<int> docs)
      : offsets_(std::move(offsets)), docs_(std::move(docs)) {
    if (offsets_.empty() || (size_t)offsets_.back() != docs_.size())
      throw std::invalid_argument("InvertedIndex: offsets do not match the posting count");
    d_offsets_.assign(offsets_);
    d_docs_.assign(docs_);
  }
  int get_num_nodes() const override { return (int)offsets_.size() - 1; }
  int get_num_elements() const override { return (int)docs_.size(); }
  const std::vector<int>& get_host_offsets() const override { return offsets_; }
  const std::vector<int>& get_host_elements() const override { return docs_; }
  int* get_device_offsets() const override { return d_offsets_.data(); }
  int* get_device_elements() const override { return d_docs_.data(); }

 private:
  std::vector<int> offsets_, docs_;
  DeviceArray<int> d_offsets_, d_docs_;
};

// ------------------------------------------------------------------ tables
// baseline/common/hash_table.cuh:12-95.  Builds both layouts; owns the
// device copy of the CSR the tables were built over.
class HashTableBuilder {
 public:
  HashTableBuilder() = default;
  virtual ~HashTableBuilder() { free_hash_tables(); }
  HashTableBuilder(const HashTableBuilder&) = delete;
  HashTableBuilder& operator=(const HashTableBuilder&) = delete;

  // hash_table.cu:69-192.  When vertexs / vertex_csr_cols are given they are
  // the (row, element) entries to insert; they must list exactly the CSR's
  // entries in row order (the case of every reference caller).
  void build_hash_tables(int num_nodes, const std::vector<int>& h_offsets,
                         const std::vector<int>& h_elements, float load_factor, int bucket_size,
                         const std::vector<int>& vertexs = {},
                         const std::vector<int>& vertex_csr_cols = {}) {
    free_hash_tables();
    if ((int64_t)h_offsets.size() != (int64_t)num_nodes + 1)
      throw std::invalid_argument("build_hash_tables: offsets must hold num_nodes + 1 entries");
    const std::vector<int>* elems = &h_elements;
    if (!vertexs.empty() || !vertex_csr_cols.empty()) {
      if (vertexs.size() != vertex_csr_cols.size() ||
          vertex_csr_cols.size() != (size_t)h_offsets.back())
        throw std::invalid_argument("build_hash_tables: entry lists do not match the offsets");
      for (int r = 0; r < num_nodes; ++r)
        for (int e = h_offsets[r]; e < h_offsets[r + 1]; ++e)
          if (vertexs[e] != r)
            throw std::invalid_argument("build_hash_tables: entries are not in CSR row order");
      elems = &vertex_csr_cols;
    }
    d_offsets_.assign(h_offsets);
    d_elements_.assign(*elems);
    check(rs_tables_build(d_offsets_.data(), num_nodes, d_elements_.data(), load_factor,
                          bucket_size, RS_BUILD_BOTH, nullptr, &tables_),
          "build_hash_tables");
    check(rs_tables_describe(tables_, &desc_), "build_hash_tables");
  }
  void free_hash_tables() {
    if (tc_plan_) rs_tc_plan_free(tc_plan_);
    tc_plan_ = nullptr;
    if (tables_) rs_tables_free(tables_);
    tables_ = nullptr;
    desc_ = rs_tables_desc{};
    d_offsets_.reset();
    d_elements_.reset();
  }
  int* get_d_hash_hierarchical() const { return desc_.d_table_hierarchical; }
  int* get_d_hash_normal() const { return desc_.d_table_normal; }
  int* get_d_hash_length() const { return desc_.d_lengths; }
  long long* get_d_hash_tables_offset() const {
    return reinterpret_cast<long long*>(desc_.d_table_offsets);
  }
  int get_max_length() const { return desc_.max_length; }
  long long get_bucket_num() const { return desc_.bucket_num; }
  int get_bucket_size() const { return desc_.bucket_size; }
  rs_tables* handle() const { return tables_; }

 protected:
  void require_built(const char* who) const {
    if (!tables_) throw Error(RS_EINVAL, std::string(who) + ": build_hash_tables first");
  }
  // The probe order of the hierarchical table (extract_hashtable_to_csr,
  // baseline/common/utils.cu:148-175).
  DeviceArray<int> probe_order() const {
    DeviceArray<int> out(d_elements_.size());
    check(rs_tables_probe_order(tables_, RS_HIERARCHICAL, out.data(), nullptr), "probe_order");
    return out;
  }
  rs_tc_plan* tc_plan() {
    if (!tc_plan_) check(rs_tc_plan_create(tables_, nullptr, &tc_plan_), "tc plan");
    return tc_plan_;
  }

  rs_tables* tables_ = nullptr;
  rs_tables_desc desc_{};
  rs_tc_plan* tc_plan_ = nullptr;
  DeviceArray<int> d_offsets_, d_elements_;
};

// Times `enqueue` on the default stream with HIP events; ms.
template <typename F>
inline float time_ms(F&& enqueue) {
  hipEvent_t a, b;
  check_hip(hipEventCreate(&a), "hipEventCreate");
  check_hip(hipEventCreate(&b), "hipEventCreate");
  check_hip(hipEventRecord(a, nullptr), "hipEventRecord");
  enqueue();
  check_hip(hipEventRecord(b, nullptr), "hipEventRecord");
  check_hip(hipEventSynchronize(b), "hipEventSynchronize");
  float ms = 0.f;
  check_hip(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms;
}

// ------------------------------------------------------------------ TC
// baseline/tc/tc.cuh:29-37 over an oriented CSR (tables built over it).
class TCBaseline : public HashTableBuilder {
 public:
  void prepare_vertex_list(const CSRGraph&) {}
  void allocate_buffers() {
    if (!d_total_count_.size()) d_total_count_.resize(1);
  }
  // Probe keys in the hierarchical table's (bucket, slot) order, outside the
  // timed region (tc.cuh:46).
  void reorder_csr_by_hash_layout(CSRGraph&, int) {
    require_built("reorder_csr_by_hash_layout");
    sorted_keys_ = probe_order();
  }
  std::pair<unsigned long long, float> run_hierarchical(CSRGraph& graph, int CHUNK_SIZE,
                                                        int grid_size, int block_size,
                                                        int bucket_size, bool sorted) {
    return run(graph, RS_HIERARCHICAL, bucket_size, sorted, CHUNK_SIZE, grid_size, block_size);
  }
  std::pair<unsigned long long, float> run_normal(CSRGraph& graph, int CHUNK_SIZE, int grid_size,
                                                  int block_size, int bucket_size, bool sorted) {
    return run(graph, RS_NORMAL, bucket_size, sorted, CHUNK_SIZE, grid_size, block_size);
  }
  unsigned long long get_total_count() const { return total_; }

 private:
  std::pair<unsigned long long, float> run(CSRGraph& graph, int layout, int bucket_size,
                                           bool sorted, int, int, int) {
    require_built("TCBaseline::run");
    if (bucket_size != desc_.bucket_size)
      throw Error(RS_EINVAL, "TCBaseline::run: bucket_size differs from the built tables'");
    if (graph.get_device_offsets() && graph.get_num_elements() != (int)desc_.num_elements)
      throw Error(RS_EINVAL, "TCBaseline::run: graph differs from the built tables'");
    allocate_buffers();
    if (sorted && layout == RS_HIERARCHICAL && !sorted_keys_.size()) sorted_keys_ = probe_order();
    const int32_t* keys = sorted && layout == RS_HIERARCHICAL ? sorted_keys_.data() : nullptr;
    rs_tc_plan* plan = tc_plan();
    const float ms = time_ms([&] {
      check(rs_tc_count_async(tables_, plan, layout, keys,
                              reinterpret_cast<uint64_t*>(d_total_count_.data()), nullptr),
            "rs_tc_count_async");
    });
    check(rs_tables_status(tables_, nullptr), "TCBaseline::run");
    total_ = d_total_count_.to_host()[0];
    return {total_, ms};
  }
  DeviceArray<unsigned long long> d_total_count_;
  DeviceArray<int> sorted_keys_;
  unsigned long long total_ = 0;
};

// Float threshold -> the fraction p/q (q <= 4096) whose float32 value it is.
inline std::pair<int, int> threshold_fraction(float threshold) {
  for (int q = 1; q <= 4096; ++q) {
    const long p = std::lround((double)threshold * q);
    if (p >= 0 && p <= q && (float)((double)p / q) == threshold) return {(int)p, q};
  }
  throw Error(RS_EINVAL, "threshold is not a fraction p/q with q <= 4096");
}

// ------------------------------------------------------------------ SSS
// baseline/sss/sss.cuh:26-35: candidate pairs (vertexs[e], csr_cols[e]).
class SSSBaseline : public HashTableBuilder {
 public:
  void load_vertex_pairs(const std::string& vertexs_path) { vertexs_ = read_i32_vec(vertexs_path); }
  void set_vertex_pairs(std::vector<int> vertexs) { vertexs_ = std::move(vertexs); }
  void allocate_buffers() {}
  void reorder_csr_by_hash_layout(CSRGraph&, int) {
    require_built("reorder_csr_by_hash_layout");
    sorted_keys_ = probe_order();
  }
  void pre_sort_csr_cols(CSRGraph& graph, int bucket_size) {
    reorder_csr_by_hash_layout(graph, bucket_size);
  }
  std::pair<int, float> run_hierarchical(CSRGraph& graph, int CHUNK_SIZE, int grid_size,
                                         int block_size, int bucket_size, float threshold,
                                         bool sorted) {
    return run(graph, RS_HIERARCHICAL, bucket_size, threshold, sorted);
  }
  std::pair<int, float> run_normal(CSRGraph& graph, int CHUNK_SIZE, int grid_size, int block_size,
                                   int bucket_size, float threshold, bool sorted) {
    return run(graph, RS_NORMAL, bucket_size, threshold, sorted);
  }
  int get_result_count() const { return count_; }
  int get_num_pairs() const { return (int)vertexs_.size(); }

 private:
  std::pair<int, float> run(CSRGraph& graph, int layout, int bucket_size, float threshold,
                            bool sorted) {
    require_built("SSSBaseline::run");
    if (bucket_size != desc_.bucket_size)
      throw Error(RS_EINVAL, "SSSBaseline::run: bucket_size differs from the built tables'");
    const std::vector<int>& cols = graph.get_host_elements();
    if (vertexs_.size() != cols.size())
      throw Error(RS_EINVAL, "SSSBaseline::run: one vertex per CSR column expected");
    const auto pq = threshold_fraction(threshold);
    const int64_t P = (int64_t)vertexs_.size();
    if (d_u_.size() != vertexs_.size()) d_u_.assign(vertexs_);
    DeviceArray<int> flags(P), sizes(P);
    if (sorted && layout == RS_HIERARCHICAL && !sorted_keys_.size()) sorted_keys_ = probe_order();
    const int32_t* keys = sorted && layout == RS_HIERARCHICAL ? sorted_keys_.data() : nullptr;
    const float ms = time_ms([&] {
      check(rs_sss_verify_async(tables_, layout, keys, d_u_.data(), graph.get_device_elements(), P,
                                pq.first, pq.second, flags.data(), sizes.data(), nullptr),
            "rs_sss_verify_async");
    });
    check(rs_tables_status(tables_, nullptr), "SSSBaseline::run");
    int64_t c = 0;
    for (int f : flags.to_host()) c += f;
    count_ = (int)c;
    return {count_, ms};
  }
  std::vector<int> vertexs_;
  DeviceArray<int> d_u_, sorted_keys_;
  int count_ = 0;
};

// ------------------------------------------------------------------ IR
// baseline/ir/ir.cuh:39-46 over an inverted index (tables built over it).
class IRBaseline : public HashTableBuilder {
 public:
  void pre_sort_inverted_index(const InvertedIndex&, int) {
    require_built("pre_sort_inverted_index");
    sorted_keys_ = probe_order();
  }
  void load_queries(const std::string& query_path, const std::string& query_offsets_path,
                    const std::string& query_num_path) {
    set_queries(read_i32_vec(query_path), read_i32_vec(query_offsets_path),
                read_int_binary(query_num_path));
  }
  void set_queries(std::vector<int> query, std::vector<int> query_offsets, int query_num) {
    if ((int)query_offsets.size() != query_num + 1 || (size_t)query_offsets.back() != query.size())
      throw Error(RS_EINVAL, "IRBaseline: query offsets do not match the query list");
    query_ = std::move(query);
    query_offsets_ = std::move(query_offsets);
    query_num_ = query_num;
  }
  // Result regions of |list(term0)| entries per query (ir.cu:324-338).
  void allocate_result_buffers(const InvertedIndex& index) {
    const std::vector<int>& off = index.get_host_offsets();
    std::vector<int64_t> roff(query_num_ + 1, 0);
    for (int q = 0; q < query_num_; ++q) {
      const int a = query_offsets_[q], z = query_offsets_[q + 1];
      const int64_t len = a < z ? off[query_[a] + 1] - off[query_[a]] : 0;
      roff[q + 1] = roff[q] + len;
    }
    d_query_.assign(query_);
    d_query_offsets_.assign(query_offsets_);
    d_result_offsets_.assign(roff);
    d_result_.resize((size_t)roff.back());
    d_workspace_.resize((size_t)roff.back());
    d_result_count_.resize(query_num_);
  }
  std::pair<int, float> run_hierarchical(int CHUNK_SIZE, int grid_size, int block_size,
                                         int bucket_size, bool sorted) {
    return run(RS_HIERARCHICAL, bucket_size, sorted);
  }
  std::pair<int, float> run_normal(int CHUNK_SIZE, int grid_size, int block_size, int bucket_size,
                                   bool sorted) {
    return run(RS_NORMAL, bucket_size, sorted);
  }
  std::vector<int> get_results() const { return d_result_count_.to_host(); }
  // Documents of query q (the first get_results()[q] entries of its region).
  std::vector<int> get_result_docs(int q) const {
    const std::vector<int64_t> roff = d_result_offsets_.to_host();
    const std::vector<int> cnt = d_result_count_.to_host();
    std::vector<int> out((size_t)cnt[q]);
    if (cnt[q])
      check_hip(hipMemcpy(out.data(), d_result_.data() + roff[q], out.size() * sizeof(int),
                          hipMemcpyDeviceToHost),
                "hipMemcpy");
    return out;
  }
  int get_query_num() const { return query_num_; }
  // Device buffers (baseline/ir/ir.cuh:55-62), valid after
  // allocate_result_buffers; the reference's work counter is not needed.
  int* get_d_query() const { return d_query_.data(); }
  int* get_d_query_offsets() const { return d_query_offsets_.data(); }
  int* get_d_result() const { return d_result_.data(); }
  long long* get_d_result_offsets() const {
    return reinterpret_cast<long long*>(d_result_offsets_.data());
  }
  int* get_d_result_count() const { return d_result_count_.data(); }
  int* get_d_G_index() const { return nullptr; }

 private:
  std::pair<int, float> run(int layout, int bucket_size, bool sorted) {
    require_built("IRBaseline::run");
    if (bucket_size != desc_.bucket_size)
      throw Error(RS_EINVAL, "IRBaseline::run: bucket_size differs from the built tables'");
    if (d_result_count_.size() != (size_t)query_num_)
      throw Error(RS_EINVAL, "IRBaseline::run: allocate_result_buffers first");
    if (sorted && layout == RS_HIERARCHICAL && !sorted_keys_.size()) sorted_keys_ = probe_order();
    const int32_t* term0 = sorted && layout == RS_HIERARCHICAL ? sorted_keys_.data() : nullptr;
    const float ms = time_ms([&] {
      check(rs_ir_query_async(tables_, layout, term0, d_query_.data(), d_query_offsets_.data(),
                              query_num_, d_result_offsets_.data(), d_result_.data(),
                              d_result_count_.data(), d_workspace_.data(), nullptr),
            "rs_ir_query_async");
    });
    check(rs_tables_status(tables_, nullptr), "IRBaseline::run");
    int64_t s = 0;
    for (int c : get_results()) s += c;
    return {(int)s, ms};
  }
  std::vector<int> query_, query_offsets_;
  int query_num_ = 0;
  DeviceArray<int> d_query_, d_query_offsets_, d_result_, d_workspace_, d_result_count_;
  DeviceArray<int64_t> d_result_offsets_;
  DeviceArray<int> sorted_keys_;
};

// ------------------------------------------------------------ competitor tables
// The free functions of baseline/tc/tc_cuco.cuh:1-32, baseline/sss/
// sss_baselines.h:14-70 and sss_cuco.cuh over the competitor tables
// (rs_comp_*): the tables are built from the device CSR (which must hold the
// host CSR's values), then the probe kernel alone is timed.  grid_size,
// block_size and CHUNK_SIZE are ignored; the threshold is read as its exact
// fraction; keys a table could not place are reported on stderr, as the
// reference warns of them.
namespace detail {
class CompHandle {
 public:
  CompHandle(const int* d_offsets, int num_sets, const int* d_elements, int kind, float load_factor) {
    check(rs_comp_build(d_offsets, num_sets, d_elements, kind, load_factor, nullptr, &c_),
          "rs_comp_build");
    int64_t slots = 0, failed = 0;
    check(rs_comp_info(c_, &slots, &failed), "rs_comp_info");
    if (failed) std::fprintf(stderr, "warning: %lld keys could not be inserted\n", (long long)failed);
  }
  ~CompHandle() { rs_comp_free(c_); }
  CompHandle(const CompHandle&) = delete;
  CompHandle& operator=(const CompHandle&) = delete;
  rs_comp* get() const { return c_; }

 private:
  rs_comp* c_ = nullptr;
};

inline std::pair<int, float> run_sss_kind(int kind, int num_pairs, int num_nodes, int const* d_vertexs,
                                          int const* d_csr_cols_for_vertexs, int const* d_csr_cols,
                                          int const* d_csr_offsets, float threshold, float load_factor) {
  const auto pq = threshold_fraction(threshold);
  CompHandle c(d_csr_offsets, num_nodes, d_csr_cols, kind, load_factor);
  DeviceArray<int> flags(num_pairs), sizes(num_pairs);
  const float ms = time_ms([&] {
    check(rs_comp_sss_verify_async(c.get(), d_vertexs, d_csr_cols_for_vertexs, num_pairs, pq.first,
                                   pq.second, flags.data(), sizes.data(), nullptr),
          "rs_comp_sss_verify_async");
  });
  check(rs_comp_status(c.get(), nullptr), "rs_comp_sss_verify_async");
  int64_t n = 0;
  for (int f : flags.to_host()) n += f;
  return {(int)n, ms};
}
}  // namespace detail

// tc_cuco.cuh: TC through linear-probing (cuCollections-style) tables.
inline std::pair<unsigned long long, float> run_tc_cuco(
    int num_nodes, int num_edges, int const* d_vertexs, int const* d_csr_row,
    int const* d_csr_cols_for_traversal, std::vector<int> const& csr_row_host,
    std::vector<int> const& csr_cols_host, int grid_size, int block_size, int CHUNK_SIZE,
    float load_factor, hipStream_t stream = nullptr) {
  (void)num_edges, (void)d_vertexs, (void)csr_row_host, (void)csr_cols_host, (void)grid_size,
      (void)block_size, (void)CHUNK_SIZE, (void)stream;
  detail::CompHandle c(d_csr_row, num_nodes, d_csr_cols_for_traversal, RS_COMP_LINEAR, load_factor);
  uint64_t count = 0;
  check(rs_comp_tc_count(c.get(), &count, nullptr), "rs_comp_tc_count");  // warm-up (row ids)
  const float ms = time_ms([&] { check(rs_comp_tc_count(c.get(), &count, nullptr), "rs_comp_tc_count"); });
  return {(unsigned long long)count, ms};
}

// ir_cuco.cuh: conjunctive queries through linear-probing tables built over
// the inverted index; returns (total result count, kernel ms).
inline std::pair<int, float> run_ir_cuco(
    int inverted_index_num, int query_num, int const* d_inverted_index,
    int const* d_inverted_index_offsets, int const* d_query, int const* d_query_offsets,
    int* d_result, long long const* d_result_offsets, int* d_result_count, int* d_G_index,
    int CHUNK_SIZE, std::vector<int> const& inverted_index_offsets_host,
    std::vector<int> const& inverted_index_host, int grid_size, int block_size, float load_factor,
    hipStream_t stream = nullptr) {
  (void)d_G_index, (void)CHUNK_SIZE, (void)inverted_index_offsets_host, (void)inverted_index_host,
      (void)grid_size, (void)block_size, (void)stream;
  static_assert(sizeof(long long) == sizeof(int64_t), "64-bit result offsets");
  detail::CompHandle c(d_inverted_index_offsets, inverted_index_num, d_inverted_index, RS_COMP_LINEAR,
                       load_factor);
  const float ms = time_ms([&] {
    check(rs_comp_ir_query_async(c.get(), d_query, d_query_offsets, query_num,
                                 reinterpret_cast<const int64_t*>(d_result_offsets), d_result,
                                 d_result_count, nullptr),
          "rs_comp_ir_query_async");
  });
  check(rs_comp_status(c.get(), nullptr), "rs_comp_ir_query_async");
  std::vector<int> counts(query_num);
  if (query_num)
    check_hip(hipMemcpy(counts.data(), d_result_count, sizeof(int) * query_num, hipMemcpyDeviceToHost),
              "hipMemcpy");
  int64_t total = 0;
  for (int v : counts) total += v;
  return {(int)total, ms};
}

#define RESET_AMD_SSS_COMP(name, kind)                                                             \
  inline std::pair<int, float> name(                                                               \
      int num_pairs, int num_nodes, int const* d_vertexs, int const* d_csr_cols_for_vertexs,       \
      int const* d_csr_cols, int const* d_csr_offsets, std::vector<int> const& csr_offsets_host,   \
      std::vector<int> const& csr_cols_host, float threshold, int grid_size, int block_size,       \
      int CHUNK_SIZE, float load_factor = 0.2f, hipStream_t stream = nullptr) {                    \
    (void)csr_offsets_host, (void)csr_cols_host, (void)grid_size, (void)block_size,                \
        (void)CHUNK_SIZE, (void)stream;                                                            \
    return detail::run_sss_kind(kind, num_pairs, num_nodes, d_vertexs, d_csr_cols_for_vertexs,     \
                                d_csr_cols, d_csr_offsets, threshold, load_factor);                \
  }
RESET_AMD_SSS_COMP(run_sss_cuco, RS_COMP_LINEAR)          // sss_cuco.cuh
RESET_AMD_SSS_COMP(run_sss_cuckoo, RS_COMP_CUCKOO)        // sss_baselines.h:32-38
RESET_AMD_SSS_COMP(run_sss_hopscotch, RS_COMP_HOPSCOTCH)  // sss_baselines.h:43-49
RESET_AMD_SSS_COMP(run_sss_roaring, RS_COMP_ROARING)      // sss_baselines.h:54-60
#undef RESET_AMD_SSS_COMP

// ------------------------------------------------------------------ pairs
// set_size/intersection.cu:128-181 for one pair, and the batched form for
// many: |A_i cap B_i| for pairs given as host vectors.
inline std::vector<int> intersect_pairs(const std::vector<int64_t>& a_off,
                                        const std::vector<int>& a,
                                        const std::vector<int64_t>& b_off,
                                        const std::vector<int>& b, float* kernel_ms = nullptr) {
  const int64_t P = (int64_t)a_off.size() - 1;
  if (P < 0 || (int64_t)b_off.size() != P + 1)
    throw Error(RS_EINVAL, "intersect_pairs: offset arrays must have P + 1 entries");
  DeviceArray<int64_t> da_off(a_off), db_off(b_off);
  DeviceArray<int> da(a), db(b), sizes(P);
  rs_pairs_opts o;
  rs_pairs_default_opts(&o);
  rs_pairs_plan* plan = nullptr;
  check(rs_pairs_plan_create(P, &o, &plan), "rs_pairs_plan_create");
  try {
    const float ms = time_ms([&] {
      check(rs_pairs_intersect_batched_async(plan, da_off.data(), da.data(), db_off.data(),
                                             db.data(), P, sizes.data(), nullptr),
            "rs_pairs_intersect_batched_async");
    });
    check(rs_pairs_plan_status(plan, nullptr), "intersect_pairs");
    if (kernel_ms) *kernel_ms = ms;
  } catch (...) {
    rs_pairs_plan_free(plan);
    throw;
  }
  rs_pairs_plan_free(plan);
  return sizes.to_host();
}

}  // namespace reset_amd

#endif  // RESET_AMD_HPP
