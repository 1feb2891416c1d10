// A C++ caller of include/cytom.h written the way a maintainer binds the
// reference's execute<VertexType, Algorithm, is_weighted> body
// (src/include/main/execute.h:534-771) to the engine: the binary edge file is
// read as file_edge_t or edge_compact_weighted_t records (util/datatypes.h:
// 41-44, 60-64), every batch_insertion_count records form one update batch
// (EdgeInserter::run, edge-inserter.h:53-194), the program runs after every
// batch (edge-inserter.h:129-184), and the values are written in the
// ResultsWriter layout (results-writer.h:54-67).
//
// Usage: execute_shim <edges.bin> <record_bytes 8|12> <bfs|cc|sssp|pagerank-delta>
//                     <n_vertices> <batch_insertion_count> <algo_start_id> <results.bin>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "cytom.h"

namespace {

struct file_edge_t {  // util/datatypes.h:41-44
  int32_t src;
  int32_t tgt;
};
struct edge_compact_weighted_t {  // util/datatypes.h:60-64
  uint32_t src;
  uint32_t tgt;
  float weight;
};
struct PageRankDeltaVertexType_t {  // pagerank-delta.cc:30-34
  float delta_sum;
  float rank;
  float delta;
};

// The ConfigContext fields the binding reads (execute.h:45-362).
struct ShimContext {
  uint64_t max_count_vertices;
  uint32_t algorithm_iterations;
  double delta;
  uint32_t algo_start_id;
  uint64_t batch_insertion_count;
};

template <typename VertexType, typename Record>
int execute_mi355x(const ShimContext& c, uint32_t kind, const Record* records, uint64_t count_edges,
                   std::vector<VertexType>& out, uint64_t* total_iterations) {
  cytom_config cfg{};
  cfg.n_vertices = c.max_count_vertices;  // normalizeConfig, config.cc:11-19
  cfg.vertices_per_tile = 0;              // engine default
  cfg.max_iterations = c.algorithm_iterations;
  cfg.delta = c.delta;
  cfg.weighted = kind == CYTOM_SSSP;
  cytom_ctx* ctx = nullptr;
  if (cytom_create(&cfg, &ctx)) return -1;
  // The Algorithm's statics, checked by the engine at registration.
  cytom_program_desc d;
  if (cytom_program_statics(kind, &d)) return -1;
  d.source = c.algo_start_id;
  int prog = -1;
  int rc = cytom_program_register(ctx, &d, &prog);
  if (rc) {
    std::fprintf(stderr, "register: %s\n", cytom_last_error(ctx));
    cytom_destroy(ctx);
    return rc;
  }
  *total_iterations = 0;
  for (uint64_t b = 0, n = 0; b < count_edges; b += n) {
    n = std::min<uint64_t>(c.batch_insertion_count, count_edges - b);
    cytom_batch_stats bs;
    rc = cytom_update_batch_records(ctx, records + b, sizeof(Record), n, 1, nullptr, 0, &bs);
    if (rc < 0) {
      std::fprintf(stderr, "update: %s\n", cytom_last_error(ctx));
      break;
    }
    cytom_run_stats rs;
    if ((rc = cytom_execute(ctx, prog, &rs)) < 0) {
      std::fprintf(stderr, "execute: %s\n", cytom_last_error(ctx));
      break;
    }
    *total_iterations += rs.iterations;
  }
  uint64_t V = 0;
  cytom_geometry(ctx, &V, nullptr, nullptr);
  out.resize(V);
  if (rc >= 0 && cytom_read_results(ctx, prog, out.data(), out.size() * sizeof(VertexType)) < 0)
    rc = -1;
  cytom_destroy(ctx);
  return rc < 0 ? rc : 0;
}

template <typename Record>
std::vector<Record> read_records(const char* path) {
  std::vector<Record> r;
  FILE* f = std::fopen(path, "rb");
  if (!f) return r;
  std::fseek(f, 0, SEEK_END);
  const long bytes = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  r.resize((size_t)bytes / sizeof(Record));
  if (std::fread(r.data(), sizeof(Record), r.size(), f) != r.size()) r.clear();
  std::fclose(f);
  return r;
}

template <typename VertexType, typename Record>
int run(const ShimContext& c, uint32_t kind, const char* edges, const char* results) {
  const std::vector<Record> rec = read_records<Record>(edges);
  std::vector<VertexType> out;
  uint64_t iters = 0;
  const int rc = execute_mi355x<VertexType>(c, kind, rec.data(), rec.size(), out, &iters);
  if (rc) return 1;
  FILE* f = std::fopen(results, "wb");
  if (!f || std::fwrite(out.data(), sizeof(VertexType), out.size(), f) != out.size()) return 1;
  std::fclose(f);
  std::printf("{\"edges\": %zu, \"vertices\": %zu, \"iterations\": %llu}\n", rec.size(),
              out.size(), (unsigned long long)iters);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 8) {
    std::fprintf(stderr,
                 "usage: %s <edges.bin> <record_bytes> <kind> <n_vertices> <batch> <source> "
                 "<results.bin>\n",
                 argv[0]);
    return 2;
  }
  const uint32_t rb = (uint32_t)std::atoi(argv[2]);
  const std::string name = argv[3];
  ShimContext c{};
  c.max_count_vertices = std::strtoull(argv[4], nullptr, 10);
  c.algorithm_iterations = 100;  // run.py:119
  c.delta = 0.01;                // pagerank-delta.cc:27 (DELTA)
  c.batch_insertion_count = std::strtoull(argv[5], nullptr, 10);
  c.algo_start_id = (uint32_t)std::strtoul(argv[6], nullptr, 10);
  if (name == "bfs" && rb == 8) return run<uint32_t, file_edge_t>(c, CYTOM_BFS, argv[1], argv[7]);
  if (name == "cc" && rb == 8) return run<uint32_t, file_edge_t>(c, CYTOM_CC, argv[1], argv[7]);
  if (name == "sssp" && rb == 12)
    return run<float, edge_compact_weighted_t>(c, CYTOM_SSSP, argv[1], argv[7]);
  if (name == "pagerank-delta" && rb == 8)
    return run<PageRankDeltaVertexType_t, file_edge_t>(c, CYTOM_PRD, argv[1], argv[7]);
  std::fprintf(stderr, "unsupported kind/record size: %s/%u\n", name.c_str(), rb);
  return 2;
}
