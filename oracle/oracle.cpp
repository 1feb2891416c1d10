// CPU oracle for the evolving-graph hot path -- TEST INFRASTRUCTURE ONLY.
//
// A plain C++ restatement of the reference's streaming ingest and
// incremental executors under a stated semantics contract (each deviation
// from the literal reference code is a named switch below), and
// textbook checkers.  It is the checker for the HIP library and the CPU
// baseline timed beside it; it is never part of the product path.
//
// Contract switches relative to the literal reference code:
//   D1  RMAT thresholds are initialised (rmat-context.h:14-15 never sets
//       a_/ab_/abc_); a = 0.57f, b = c = 0.19f as promoted to double
//       (datatypes.h:17-19, rmat-edge-provider.h:37).
//   D2  SSSP keeps u32 distances, u32 weights >= 1, delta 0, source set to 0
//       (sssp.cc:90-99,171-195 with the bfs.cc:193-196 source fix).
//   D3  The active set at the start of a run is init activations, hook
//       activations and leftovers of an unconverged run
//       (edge-inserter.h:149-151, continuous-algorithm-executor.h:179-216).
//   D4  Tile activity and apply gating use every bit of the bitmap
//       (tile-manager.h:413-418, apply-executor.h:69-82).
//   D5  Deletion removes one live copy of the named edge and decrements both
//       degrees (tile-manager.h:165-190, meta-tile-manager.h:127-140).
//   D6  Hooks run in stream order on one thread (meta-tile-manager.h:192-198).
//   D11 Deletion re-execution for BFS/CC/SSSP uses the exact "tight edge"
//       test against the last converged values instead of
//       critical_neighbor (bfs.cc:132-137), which goes stale after
//       hook-driven improvements.
//   D12 A batch is applied as a unit: all inserts, then all deletions, then
//       the hooks of the inserted edges that are still live, in stream
//       order, reading the degrees after the whole batch.
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <limits>
#include <mutex>
#include <queue>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

constexpr uint32_t INF = UINT32_MAX;
constexpr double ALPHA = 0.85;  // pagerank-delta.cc:19
constexpr uint32_t TILE_BATCHING = 256;  // config.py default tile batch
constexpr uint64_t APPLY_STEP = 4096;    // apply-executor.h:17

// ---------------------------------------------------------------- hashing
// util.h:178-186 (32-bit) and util.h:188-196 (64-bit), Ligra's hash.
inline uint32_t hash32(uint32_t a) {
  a = (a + 0x7ed55d16u) + (a << 12);
  a = (a ^ 0xc761c23cu) ^ (a >> 19);
  a = (a + 0x165667b1u) + (a << 5);
  a = (a + 0xd3a2646cu) ^ (a << 9);
  a = (a + 0xfd7046c5u) + (a << 3);
  a = (a ^ 0xb55a4f09u) ^ (a >> 16);
  return a;
}

inline uint64_t hash64(uint64_t a) {
  a = (a + 0x7ed55d166bef7a1dull) + (a << 12);
  a = (a ^ 0xc761c23c510fa2ddull) ^ (a >> 9);
  a = (a + 0x165667b183a9c0e1ull) + (a << 59);
  a = (a + 0xd3a2646cab3487e3ull) ^ (a << 49);
  a = (a + 0xfd7046c5ef9ab54cull) + (a << 3);
  a = (a ^ 0xb55a4f090dd4a67bull) ^ (a >> 32);
  return a;
}

// util.h:198-200
inline double hash_double(uint64_t i) {
  return (double)hash64(i) / (double)UINT64_MAX;
}

// D1: thresholds as the float literals promoted to double, summed in order.
const double RMAT_A = (double)0.57f;
const double RMAT_AB = RMAT_A + (double)0.19f;
const double RMAT_ABC = RMAT_AB + (double)0.19f;

// rmat-context.h:39-65, iterative form: levels are consumed from the least
// significant bit (stride offset scale-1) up to the most significant bit
// (offset 0).
inline void rmat_iterative(uint32_t scale, uint64_t start, uint64_t stride,
                           uint32_t* s, uint32_t* d) {
  uint64_t src = 0, dst = 0;
  for (uint32_t k = 1; k <= scale; ++k) {
    uint64_t r_index = start + (uint64_t)(scale - k) * stride;
    double r = hash_double(r_index);
    uint64_t half = 1ull << (k - 1);
    if (r < RMAT_A) {
    } else if (r < RMAT_AB) {
      dst += half;
    } else if (r < RMAT_ABC) {
      src += half;
    } else {
      src += half;
      dst += half;
    }
  }
  *s = (uint32_t)src;
  *d = (uint32_t)dst;
}

// rmat-context.h:19-37, recursive form (kept separate on purpose: it pins the
// level/offset order of the iterative one).
void rmat_recursive(uint64_t n, uint64_t start, uint64_t stride, uint64_t* s,
                    uint64_t* d) {
  if (n == 1) {
    *s = 0;
    *d = 0;
    return;
  }
  uint64_t es, ed;
  rmat_recursive(n / 2, start + stride, stride, &es, &ed);
  double r = hash_double(start);
  if (r < RMAT_A) {
    *s = es;
    *d = ed;
  } else if (r < RMAT_AB) {
    *s = es;
    *d = ed + n / 2;
  } else if (r < RMAT_ABC) {
    *s = es + n / 2;
    *d = ed;
  } else {
    *s = es + n / 2;
    *d = ed + n / 2;
  }
}

// Stream constants (build-defined; no reference counterpart).
constexpr uint64_t WEIGHT_SALT = 0x7765696768747321ull;  // "weights!"
constexpr uint64_t DELETE_SALT = 0x64656c6574657321ull;  // "delete!"

// ------------------------------------------------------------ thread pool
class Pool {
 public:
  explicit Pool(int n) : n_(n < 1 ? 1 : n) {
    for (int i = 1; i < n_; ++i) workers_.emplace_back([this, i] { loop(i); });
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_ = true;
      ++gen_;
    }
    cv_.notify_all();
    for (auto& t : workers_) t.join();
  }
  int size() const { return n_; }
  // Run fn(thread_id) on every thread, wait for all.
  void run(const std::function<void(int)>& fn) {
    if (n_ == 1) {
      fn(0);
      return;
    }
    {
      std::lock_guard<std::mutex> lk(m_);
      fn_ = &fn;
      pending_ = n_ - 1;
      ++gen_;
    }
    cv_.notify_all();
    fn(0);
    std::unique_lock<std::mutex> lk(m_);
    done_cv_.wait(lk, [this] { return pending_ == 0; });
    fn_ = nullptr;
  }

 private:
  void loop(int id) {
    uint64_t seen = 0;
    for (;;) {
      const std::function<void(int)>* fn;
      {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] { return gen_ != seen; });
        seen = gen_;
        if (stop_) return;
        fn = fn_;
      }
      (*fn)(id);
      {
        std::lock_guard<std::mutex> lk(m_);
        if (--pending_ == 0) done_cv_.notify_one();
      }
    }
  }
  int n_;
  std::vector<std::thread> workers_;
  std::mutex m_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(int)>* fn_ = nullptr;
  int pending_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

// ---------------------------------------------------------------- bitmaps
// algorithm-common.h:5-12, util.h:36-66: one bit per vertex in bytes.
inline bool bit_get(const std::vector<uint8_t>& a, uint64_t i) {
  return (a[i >> 3] >> (i & 7)) & 1u;
}
inline void bit_set(std::vector<uint8_t>& a, uint64_t i) {
  a[i >> 3] |= (uint8_t)(1u << (i & 7));
}
inline void bit_set_atomic(std::vector<uint8_t>& a, uint64_t i) {
  __atomic_fetch_or(&a[i >> 3], (uint8_t)(1u << (i & 7)), __ATOMIC_RELAXED);
}

inline uint32_t sat_add(uint32_t a, uint32_t b) {
  uint64_t s = (uint64_t)a + b;
  return s >= INF ? INF : (uint32_t)s;
}

// pagerank-delta.cc:30-43
struct PRDV {
  float delta_sum;
  float rank;
  float delta;
};

struct Tile {
  std::vector<uint16_t> s, d;  // tile-local ids (datatypes.h:77-92)
  std::vector<uint32_t> w;
  std::vector<uint8_t> del;    // deleted-edges bitset (tile-manager.h:140-141)
  uint64_t live = 0;
};

struct Prog {
  int kind = 0;
  uint32_t source = 0;
  std::vector<uint32_t> teleport;  // PPR: distinct teleport vertices
  std::vector<uint32_t> cur, nxt;
  std::vector<PRDV> pcur, pnxt;
  std::vector<uint8_t> active, active_next, changed;
  bool first_iter = true;  // pagerank-delta.cc GLOBAL_FIRST_ITERATION
  bool fixpoint = false;   // values are the converged result of the live graph
  int last_reexec = 0;
};

struct DelRec {
  uint32_t src, dst, w;
};

}  // namespace

struct orc_stream {
  uint64_t hash_rmat = 0, hash_del = 0;
  uint64_t next_index = 0;
  uint64_t batch = 0;
  std::vector<uint64_t> live;
};

struct orc_graph {
  uint64_t V = 0;
  uint32_t vpt = 0, T = 0;
  bool weighted = false;
  uint32_t max_iter = 100;
  double delta = 0.0;
  std::vector<Tile> tiles;  // tile id t = y*T + x (row_first.cc:6-11)
  std::vector<uint32_t> outdeg, indeg;
  std::vector<Prog> progs;
  Pool pool;
  // Per-thread local arrays for the pull phase.
  std::vector<std::vector<std::vector<uint32_t>>> loc_u32;
  std::vector<std::vector<std::vector<float>>> loc_f32;
  orc_graph(int threads) : pool(threads) {}
};

namespace {

// ------------------------------------------------------------- programs
inline bool is_min_prog(int k) { return k == ORC_BFS || k == ORC_CC || k == ORC_SSSP; }
inline bool uses_changed(int k) { return k != ORC_SSSP; }  // USE_CHANGED statics
inline bool prd_kind(int k) { return k == ORC_PRD || k == ORC_PPR; }

// init_vertices: bfs.cc:170-197, cc.cc:171-189, sssp.cc:171-195 (+D2),
// pagerank-delta.cc:246-270.
void init_program(orc_graph* g, Prog& P) {
  const uint64_t V = g->V;
  const size_t nb = (V + 7) / 8;
  P.active.assign(nb, 0);
  P.active_next.assign(nb, 0);
  P.changed.assign(nb, 0);
  P.first_iter = P.kind != ORC_PPR;
  P.fixpoint = false;
  if (P.kind == ORC_PRD) {
    P.pcur.assign(V, PRDV{0.f, 0.f, 1.f});
    P.pnxt.assign(V, PRDV{0.f, 0.f, 1.f});
    std::fill(P.active.begin(), P.active.end(), 0xFF);
  } else if (P.kind == ORC_PPR) {
    // Personalized form: rank = delta = (1-alpha)/|T| on the teleport set,
    // which starts active; everything else starts at zero.
    P.pcur.assign(V, PRDV{0.f, 0.f, 0.f});
    P.pnxt.assign(V, PRDV{0.f, 0.f, 0.f});
    const float r0 = (float)((1. - ALPHA) * (1.0 / (double)P.teleport.size()));
    for (uint32_t v : P.teleport) {
      P.pcur[v].rank = P.pnxt[v].rank = r0;
      P.pcur[v].delta = P.pnxt[v].delta = r0;
      bit_set(P.active, v);
    }
  } else if (P.kind == ORC_CC) {
    P.cur.resize(V);
    P.nxt.resize(V);
    for (uint64_t i = 0; i < V; ++i) P.cur[i] = P.nxt[i] = (uint32_t)i;
    std::fill(P.active.begin(), P.active.end(), 0xFF);
  } else {
    P.cur.assign(V, INF);
    P.nxt.assign(V, INF);
    P.cur[P.source] = 0;
    P.nxt[P.source] = 0;
    bit_set(P.active, P.source);
  }
  // Bits past V stay clear so frontier counts are exact.
  for (uint64_t i = V; i < nb * 8; ++i) {
    P.active[i >> 3] &= (uint8_t) ~(1u << (i & 7));
  }
}

inline bool atomic_min_u32(uint32_t* p, uint32_t v) {
  uint32_t old = __atomic_load_n(p, __ATOMIC_RELAXED);
  while (v < old) {
    if (__atomic_compare_exchange_n(p, &old, v, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED))
      return true;
  }
  return false;
}

// compareAndSwap on the delta_sum bits (pagerank-delta.cc:233-236 via
// arch.h:106-107).  Returns true if the stored value changed.
inline bool atomic_add_f32(float* p, float v) {
  uint32_t* pi = reinterpret_cast<uint32_t*>(p);
  uint32_t old = __atomic_load_n(pi, __ATOMIC_RELAXED);
  for (;;) {
    float of;
    std::memcpy(&of, &old, 4);
    float nf = v + of;  // reduceVertex: lhs(local) + rhs(global), pagerank-delta.cc:238-244
    uint32_t ni;
    std::memcpy(&ni, &nf, 4);
    if (ni == old) return false;
    if (of == nf) return false;  // operator!= compares delta_sum only
    if (__atomic_compare_exchange_n(pi, &old, ni, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED))
      return true;
  }
}

// Pull phase for one batch of tiles (algorithm-executor.h:401-505).
// Returns traversed edges.
uint64_t pull_batch(orc_graph* g, Prog& P, int tid, uint64_t t0, uint64_t t1,
                    const std::vector<uint8_t>& row_act) {
  const uint32_t T = g->T, vpt = g->vpt;
  const int k = P.kind;
  uint64_t te = 0;
  // getLocalNextArray: one local array per distinct tile row y in the batch.
  std::vector<uint32_t> ys;
  auto& lu = g->loc_u32[tid];
  auto& lf = g->loc_f32[tid];
  auto local_index = [&](uint32_t y) -> size_t {
    for (size_t i = 0; i < ys.size(); ++i)
      if (ys[i] == y) return i;
    size_t i = ys.size();
    ys.push_back(y);
    if (prd_kind(k)) {
      if (lf.size() <= i) lf.emplace_back();
      lf[i].assign(vpt, 0.f);  // init_vertices_local: memset 0
    } else {
      if (lu.size() <= i) lu.emplace_back();
      lu[i].assign(vpt, INF);  // init_vertices_local: UINT32_MAX / FLT_MAX
    }
    return i;
  };
  for (uint64_t t = t0; t < t1; ++t) {
    const uint32_t x = (uint32_t)(t % T), y = (uint32_t)(t / T);
    const Tile& tl = g->tiles[t];
    if (!row_act[x]) continue;  // isActive (D4: full-word test)
    if (tl.live == 0) continue;  // skip empty tiles
    const size_t li = local_index(y);
    const uint64_t xb = (uint64_t)x * vpt;
    const size_t n = tl.s.size();
    if (prd_kind(k)) {
      float* loc = lf[li].data();
      for (size_t i = 0; i < n; ++i) {
        if (tl.del[i]) continue;
        const uint64_t src = xb + tl.s[i];
        if (!bit_get(P.active, src)) continue;
        ++te;
        // pullGather: pagerank-delta.cc:127-132
        loc[tl.d[i]] = loc[tl.d[i]] + (P.pcur[src].delta / (float)g->outdeg[src]);
      }
    } else {
      uint32_t* loc = lu[li].data();
      for (size_t i = 0; i < n; ++i) {
        if (tl.del[i]) continue;
        const uint64_t src = xb + tl.s[i];
        if (!bit_get(P.active, src)) continue;
        ++te;
        const uint32_t u = P.cur[src];
        uint32_t& v = loc[tl.d[i]];
        if (u == INF) continue;
        if (k == ORC_BFS) {  // bfs.cc:83-91
          if (v > u + 1) v = u + 1;
        } else if (k == ORC_CC) {  // cc.cc:81-91
          if (u < v) v = u;
        } else {  // sssp.cc:90-99 (D2)
          const uint32_t c = sat_add(u, tl.w[i]);
          if (v > c) v = c;
        }
      }
    }
  }
  // applyTiles: fold the local arrays into next (algorithm-executor.h:431-505).
  for (size_t i = 0; i < ys.size(); ++i) {
    const uint64_t base = (uint64_t)ys[i] * vpt;
    if (prd_kind(k)) {
      const float* loc = lf[i].data();
      for (uint32_t j = 0; j < vpt; ++j) {
        if (loc[j] == 0.f) continue;  // == neutral_element
        if (atomic_add_f32(&P.pnxt[base + j].delta_sum, loc[j]))
          bit_set_atomic(P.changed, base + j);
      }
    } else {
      const uint32_t* loc = lu[i].data();
      for (uint32_t j = 0; j < vpt; ++j) {
        if (loc[j] == INF) continue;
        if (atomic_min_u32(&P.nxt[base + j], loc[j])) bit_set_atomic(P.changed, base + j);
      }
    }
  }
  return te;
}

// apply(): bfs.cc:101-113, cc.cc:101-111, sssp.cc:101-113,
// pagerank-delta.cc:142-174.  Returns true when the vertex converged.
inline bool apply_vertex(orc_graph* g, Prog& P, uint64_t id) {
  if (prd_kind(P.kind)) {
    PRDV& n = P.pnxt[id];
    const float delta_sum = n.delta_sum;
    if (P.first_iter) {
      n.delta = static_cast<float>((1. - ALPHA) + (ALPHA * delta_sum));
      n.rank += n.delta;
      n.delta -= 1.0;
    } else {
      if (delta_sum == 0.0) {
        P.pcur[id].delta = 0;
        n.delta = 0;
        return true;
      }
      n.delta = static_cast<float>(ALPHA * delta_sum);
      n.rank += n.delta;
    }
    P.pcur[id].delta = n.delta;
    n.delta_sum = 0.f;
    if (std::abs(n.delta) > g->delta * n.rank) {
      bit_set(P.active_next, id);
      return false;
    }
    return true;
  }
  const uint32_t nx = P.nxt[id];
  if (nx != P.cur[id]) {
    P.cur[id] = nx;
    bit_set(P.active_next, id);
    return false;
  }
  return true;
}

bool is_vertex_active_apply(const Prog& P, uint64_t id) {
  if (P.kind == ORC_PRD && P.first_iter) return true;  // pagerank-delta.cc:209-217
  if (!uses_changed(P.kind)) return true;              // sssp.cc: USE_CHANGED=false
  return bit_get(P.changed, id);
}

// executeAlgorithm: continuous-algorithm-executor.h:49-249.
void run_program(orc_graph* g, Prog& P, orc_stats* st) {
  const uint64_t V = g->V;
  const uint32_t T = g->T, vpt = g->vpt;
  const uint64_t n_tiles = (uint64_t)T * T;
  const int nt = g->pool.size();
  uint32_t it = 0;
  bool converged = false;
  uint64_t te_total = 0, act_total = 0;
  std::vector<uint8_t> row_act(T);
  while (it < g->max_iter) {
    // Frontier size and per-row activity (TileManager::isActive with D4).
    uint64_t act = 0;
    std::fill(row_act.begin(), row_act.end(), 0);
    for (uint64_t b = 0; b < P.active.size(); ++b) {
      const uint32_t byte = P.active[b];
      if (!byte) continue;
      act += __builtin_popcount(byte);
      for (uint32_t i = 0; i < 8; ++i)
        if ((byte >> i) & 1u) row_act[(b * 8 + i) / vpt] = 1;
    }
    act_total += act;
    // Pull phase: tile batches handed out through an atomic cursor
    // (tile_distribution ATOMICS, algorithm-executor.h:550-557).
    std::atomic<uint64_t> cursor{0};
    std::vector<uint64_t> te_thr(nt, 0);
    g->pool.run([&](int tid) {
      for (;;) {
        uint64_t t0 = cursor.fetch_add(TILE_BATCHING);
        if (t0 >= n_tiles) break;
        uint64_t t1 = std::min<uint64_t>(t0 + TILE_BATCHING, n_tiles);
        te_thr[tid] += pull_batch(g, P, tid, t0, t1, row_act);
      }
    });
    for (auto v : te_thr) te_total += v;
    // Apply phase in STEP_SIZE chunks (apply-executor.h:61-106).
    std::atomic<uint64_t> vcur{0};
    std::vector<uint8_t> conv_thr(nt, 1);
    g->pool.run([&](int tid) {
      for (;;) {
        uint64_t b = vcur.fetch_add(APPLY_STEP);
        if (b >= V) break;
        uint64_t e = std::min<uint64_t>(b + APPLY_STEP, V);
        for (uint64_t id = b; id < e; ++id) {
          if (!is_vertex_active_apply(P, id)) continue;
          if (!apply_vertex(g, P, id)) conv_thr[tid] = 0;
        }
      }
    });
    bool conv = true;
    for (auto c : conv_thr) conv = conv && c;
    ++it;
    // Reset phase (reset_vertices + swap of the active arrays, D3).
    P.first_iter = false;
    std::fill(P.changed.begin(), P.changed.end(), 0);
    P.active.swap(P.active_next);
    std::fill(P.active_next.begin(), P.active_next.end(), 0);
    if (conv) {
      converged = true;
      break;
    }
  }
  P.fixpoint = converged;
  if (st) {
    st->iterations = it;
    st->converged = converged ? 1 : 0;
    st->reexecuted = (uint32_t)P.last_reexec;
    st->traversed_edges = te_total;
    st->active_vertices = act_total;
  }
  (void)vpt;
}

// Exact deletion test against converged values (D11).
bool tight_deletion(const orc_graph* g, const Prog& P, const DelRec& e) {
  (void)g;
  const uint32_t u = P.cur[e.src], v = P.cur[e.dst];
  if (P.kind == ORC_BFS) return u != INF && v == u + 1;
  if (P.kind == ORC_SSSP) return u != INF && v == sat_add(u, e.w);
  return u == v && v != e.dst;  // CC
}

}  // namespace

extern "C" {

uint32_t orc_hash32(uint32_t a) { return hash32(a); }
uint64_t orc_hash64(uint64_t a) { return hash64(a); }
double orc_hash_double(uint64_t i) { return hash_double(i); }

// getEdge: rmat-context.h:67-73 with hash_ = hash(seed) (rmat-context.h:15).
void orc_rmat_edges(uint32_t scale, uint64_t seed, uint64_t first, uint64_t n,
                    uint32_t* src, uint32_t* dst) {
  const uint64_t h = hash64(seed);
  for (uint64_t j = 0; j < n; ++j) {
    const uint64_t index = first + j;
    const uint64_t start = hash64(2ull * index * h);
    const uint64_t stride = hash64((2ull * index + 1) * h);
    rmat_iterative(scale, start, stride, &src[j], &dst[j]);
  }
}

void orc_rmat_edges_recursive(uint32_t scale, uint64_t seed, uint64_t first, uint64_t n,
                              uint32_t* src, uint32_t* dst) {
  const uint64_t h = hash64(seed);
  for (uint64_t j = 0; j < n; ++j) {
    const uint64_t index = first + j;
    uint64_t s, d;
    rmat_recursive(1ull << scale, hash64(2ull * index * h), hash64((2ull * index + 1) * h),
                   &s, &d);
    src[j] = (uint32_t)s;
    dst[j] = (uint32_t)d;
  }
}

void orc_rmat_gather(uint32_t scale, uint64_t seed, const uint64_t* index, uint64_t n,
                     uint32_t* src, uint32_t* dst) {
  const uint64_t h = hash64(seed);
  for (uint64_t j = 0; j < n; ++j) {
    const uint64_t i = index[j];
    rmat_iterative(scale, hash64(2ull * i * h), hash64((2ull * i + 1) * h), &src[j], &dst[j]);
  }
}

void orc_stream_weights_u32_at(uint64_t seed, const uint64_t* index, uint64_t n, uint32_t* w) {
  const uint64_t hw = hash64(seed ^ WEIGHT_SALT);
  for (uint64_t j = 0; j < n; ++j) w[j] = 1u + (uint32_t)(hash64(hw ^ index[j]) % 255u);
}

void orc_stream_weights_u32(uint64_t seed, uint64_t first, uint64_t n, uint32_t* w) {
  const uint64_t hw = hash64(seed ^ WEIGHT_SALT);
  for (uint64_t j = 0; j < n; ++j) w[j] = 1u + (uint32_t)(hash64(hw ^ (first + j)) % 255u);
}

void orc_stream_weights_f32(uint64_t seed, uint64_t first, uint64_t n, float* w) {
  const uint64_t hw = hash64(seed ^ WEIGHT_SALT);
  for (uint64_t j = 0; j < n; ++j)
    w[j] = (float)((hash64(hw ^ (first + j)) >> 40) + 1) * (1.0f / 16777216.0f);
}

orc_stream* orc_stream_create(uint64_t seed) {
  auto* s = new orc_stream();
  s->hash_rmat = hash64(seed);
  s->hash_del = hash64(seed ^ DELETE_SALT);
  return s;
}
void orc_stream_destroy(orc_stream* s) { delete s; }
uint64_t orc_stream_live_count(const orc_stream* s) { return s->live.size(); }

// SplitMix64's output function (Steele, Lea and Flood, "Fast splittable
// pseudorandom number generators", OOPSLA 2014): a full-avalanche 64-bit mix.
static uint64_t splitmix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// Deletions: sequential swap-remove sampling, uniform without replacement
// over the live stream indices after this batch's inserts.  Draw j of batch b
// takes position floor(splitmix64(h ^ (b << 32 | j)) * n_live / 2^64).
int64_t orc_stream_next(orc_stream* s, uint64_t n_ins, uint64_t n_del, uint64_t* ins_first,
                        uint64_t* del_idx) {
  *ins_first = s->next_index;
  for (uint64_t j = 0; j < n_ins; ++j) s->live.push_back(s->next_index + j);
  s->next_index += n_ins;
  uint64_t m = std::min<uint64_t>(n_del, s->live.size());
  for (uint64_t j = 0; j < m; ++j) {
    uint64_t n_live = s->live.size();
    uint64_t r = (uint64_t)(((unsigned __int128)splitmix64(s->hash_del ^ ((s->batch << 32) | j)) *
                             n_live) >> 64);
    del_idx[j] = s->live[r];
    s->live[r] = s->live[n_live - 1];
    s->live.pop_back();
  }
  s->batch += 1;
  return (int64_t)m;
}

orc_graph* orc_create(uint64_t n_vertices, uint32_t vertices_per_tile, int weighted,
                      uint32_t max_iterations, double delta, int n_threads) {
  if (n_vertices == 0 || vertices_per_tile == 0 || vertices_per_tile > 65536) return nullptr;
  if (vertices_per_tile & (vertices_per_tile - 1)) return nullptr;
  // normalizeConfig: round V up to a power of two (config.cc:11-19).
  uint64_t V = 1;
  while (V < n_vertices) V <<= 1;
  auto* g = new orc_graph(n_threads);
  g->V = V;
  g->vpt = (uint32_t)std::min<uint64_t>(vertices_per_tile, V);
  g->T = (uint32_t)(V / g->vpt);
  g->weighted = weighted != 0;
  g->max_iter = max_iterations;
  g->delta = delta;
  g->tiles.resize((size_t)g->T * g->T);
  g->outdeg.assign(V, 0);
  g->indeg.assign(V, 0);
  g->loc_u32.resize(g->pool.size());
  g->loc_f32.resize(g->pool.size());
  return g;
}

void orc_destroy(orc_graph* g) { delete g; }
uint64_t orc_num_vertices(const orc_graph* g) { return g->V; }

int orc_register(orc_graph* g, int kind, uint32_t source) {
  if (kind < ORC_BFS || kind > ORC_PPR) return -22;
  if (kind == ORC_SSSP && !g->weighted) return -22;
  if (source >= g->V) return -22;
  Prog P;
  P.kind = kind;
  P.source = source;
  if (kind == ORC_PPR) P.teleport.assign(1, source);
  init_program(g, P);
  g->progs.push_back(std::move(P));
  return (int)g->progs.size() - 1;
}

int64_t orc_update(orc_graph* g, const uint32_t* src, const uint32_t* dst, const uint32_t* w,
                   uint64_t n_ins, const uint32_t* dsrc, const uint32_t* ddst,
                   const uint32_t* dw, uint64_t n_del) {
  const uint64_t V = g->V;
  const uint32_t vpt = g->vpt, T = g->T;
  for (uint64_t i = 0; i < n_ins; ++i) {
    if (src[i] >= V || dst[i] >= V) return -22;
    if (g->weighted && (w == nullptr || w[i] == 0)) return -22;
  }
  for (uint64_t i = 0; i < n_del; ++i)
    if (dsrc[i] >= V || ddst[i] >= V) return -22;
  // 1. Insertions in stream order (MetaTileManager::insertEdge,
  //    meta-tile-manager.h:112-125; TileManager::insertEdge tile-manager.h:136-163).
  std::vector<std::pair<uint32_t, uint32_t>> ins_slot(n_ins);
  for (uint64_t i = 0; i < n_ins; ++i) {
    const uint32_t x = src[i] / vpt, y = dst[i] / vpt;
    const uint32_t t = y * T + x;
    Tile& tl = g->tiles[t];
    ins_slot[i] = {t, (uint32_t)tl.s.size()};
    tl.s.push_back((uint16_t)(src[i] % vpt));
    tl.d.push_back((uint16_t)(dst[i] % vpt));
    if (g->weighted) tl.w.push_back(w[i]);
    tl.del.push_back(0);
    tl.live += 1;
    g->outdeg[src[i]] += 1;
    g->indeg[dst[i]] += 1;
  }
  // 2. Deletions (D5): first live copy in overflow order, per tile.
  std::vector<DelRec> found;
  if (n_del > 0) {
    std::unordered_map<uint32_t, std::vector<uint64_t>> by_tile;
    for (uint64_t i = 0; i < n_del; ++i)
      by_tile[(ddst[i] / vpt) * T + dsrc[i] / vpt].push_back(i);
    std::vector<std::pair<uint32_t, std::vector<uint64_t>>> work(by_tile.begin(), by_tile.end());
    std::vector<int32_t> hit_slot(n_del, -1);
    std::vector<uint32_t> hit_w(n_del, 0);
    std::atomic<size_t> cur{0};
    const bool match_w = g->weighted && dw != nullptr;
    g->pool.run([&](int) {
      for (;;) {
        size_t k = cur.fetch_add(1);
        if (k >= work.size()) break;
        const uint32_t t = work[k].first;
        Tile& tl = g->tiles[t];
        std::unordered_map<uint64_t, std::vector<uint64_t>> want;
        auto key_of = [&](uint32_t s, uint32_t d, uint32_t ww) {
          uint64_t key = ((uint64_t)s << 16) | d;
          if (match_w) key |= (uint64_t)ww << 32;
          return key;
        };
        for (uint64_t di : work[k].second)
          want[key_of(dsrc[di] % vpt, ddst[di] % vpt, match_w ? dw[di] : 0)].push_back(di);
        std::unordered_map<uint64_t, size_t> taken;
        size_t remaining = work[k].second.size();
        for (size_t i = 0; i < tl.s.size() && remaining > 0; ++i) {
          if (tl.del[i]) continue;
          auto it = want.find(key_of(tl.s[i], tl.d[i], match_w ? tl.w[i] : 0));
          if (it == want.end()) continue;
          size_t& nt = taken[it->first];
          if (nt >= it->second.size()) continue;
          const uint64_t di = it->second[nt++];
          tl.del[i] = 1;
          tl.live -= 1;
          hit_slot[di] = (int32_t)i;
          hit_w[di] = g->weighted ? tl.w[i] : 0;
          --remaining;
        }
      }
    });
    for (uint64_t i = 0; i < n_del; ++i) {
      if (hit_slot[i] < 0) continue;
      found.push_back(DelRec{dsrc[i], ddst[i], hit_w[i]});
    }
  }
  // 3. Per program: re-execution decision, then hooks (edgeChanged).
  for (Prog& P : g->progs) {
    bool reexec = false;
    if (P.kind == ORC_PPR) {
      // Personalized PageRank-delta restarts from its teleport set after
      // every update batch.
      P.last_reexec = (n_ins || !found.empty()) ? 1 : 0;
      if (P.last_reexec) init_program(g, P);
      continue;
    }
    if (!found.empty()) {
      if (is_min_prog(P.kind)) {
        if (!P.fixpoint) {
          reexec = true;
        } else {
          for (const DelRec& e : found)
            if (tight_deletion(g, P, e)) {
              reexec = true;
              break;
            }
        }
      } else {
        // pagerank-delta.cc:190-206
        for (const DelRec& e : found) {
          float d = P.pnxt[e.src].rank / (float)g->outdeg[e.src];
          if (d > g->delta * P.pnxt[e.dst].rank) {
            reexec = true;
            break;
          }
        }
      }
    }
    P.last_reexec = reexec ? 1 : 0;
    if (reexec) {
      init_program(g, P);  // initAlgorithm (edge-inserter.h:164-167)
      continue;
    }
    // INSERT hooks in stream order for slots that are still live.
    for (uint64_t i = 0; i < n_ins; ++i) {
      const Tile& tl = g->tiles[ins_slot[i].first];
      if (tl.del[ins_slot[i].second]) continue;
      const uint32_t s = src[i], d = dst[i];
      if (P.kind == ORC_BFS) {  // bfs.cc:115-131
        uint32_t sv = P.nxt[s];
        if (sv == INF) continue;
        if (sv + 1 < P.nxt[d]) {
          P.nxt[d] = P.cur[d] = sv + 1;
          bit_set(P.active, d);
          bit_set(P.changed, d);
          P.fixpoint = false;
        }
      } else if (P.kind == ORC_CC) {  // cc.cc:113-131
        if (P.nxt[s] < P.nxt[d]) {
          P.nxt[d] = P.cur[d] = P.nxt[s];
          bit_set(P.active, d);
          bit_set(P.changed, d);
          P.fixpoint = false;
        }
      } else if (P.kind == ORC_SSSP) {  // sssp.cc:115-133
        uint32_t sv = P.nxt[s];
        if (sv == INF) continue;
        uint32_t nv = sat_add(sv, w[i]);
        if (nv < P.nxt[d]) {
          P.nxt[d] = P.cur[d] = nv;
          bit_set(P.active, d);
          P.fixpoint = false;
        }
      } else {  // pagerank-delta.cc:176-189
        float dl = P.pnxt[s].rank / (float)g->outdeg[s];
        P.pnxt[d].delta_sum += dl;
        if (dl > g->delta * P.pnxt[d].rank) {
          bit_set(P.active, s);
          bit_set(P.changed, s);
        }
        P.fixpoint = false;
      }
    }
    if (P.kind == ORC_PRD) {  // DELETE hook side effect (pagerank-delta.cc:190-196)
      for (const DelRec& e : found) {
        float dl = P.pnxt[e.src].rank / (float)g->outdeg[e.src];
        P.pnxt[e.dst].delta_sum -= dl;
        P.fixpoint = false;
      }
    }
  }
  // 4. Degrees drop after the hooks: the reference evaluates edgeChanged(DELETE)
  // while the deleted edge still counts in its source's out-degree (its random
  // deletions, meta-tile-manager.h:200-228, leave the degrees unchanged,
  // tile-manager.h:175-189), so rank / out_degree never divides by zero there.
  // The degrees then follow MetaTileManager::deleteEdge (meta-tile-manager.h:127-140).
  for (const DelRec& e : found) {
    g->outdeg[e.src] -= 1;
    g->indeg[e.dst] -= 1;
  }
  return (int64_t)found.size();
}

int orc_execute(orc_graph* g, int prog, orc_stats* out) {
  if (prog < 0 || prog >= (int)g->progs.size()) return -22;
  Prog& P = g->progs[prog];
  run_program(g, P, out);
  return 0;
}

int orc_read(const orc_graph* g, int prog, void* out) {
  if (prog < 0 || prog >= (int)g->progs.size()) return -22;
  const Prog& P = g->progs[prog];
  if (prd_kind(P.kind)) {
    float* r = static_cast<float*>(out);
    for (uint64_t i = 0; i < g->V; ++i) r[i] = P.pnxt[i].rank;
  } else {
    std::memcpy(out, P.cur.data(), g->V * 4);
  }
  return 0;
}

int orc_set_sources(orc_graph* g, int prog, const uint32_t* sources, uint32_t n) {
  if (prog < 0 || prog >= (int)g->progs.size() || !sources || n == 0) return -22;
  for (uint32_t i = 0; i < n; ++i)
    if (sources[i] >= g->V) return -22;
  Prog& P = g->progs[prog];
  if (P.kind == ORC_BFS || P.kind == ORC_SSSP) {
    if (n != 1) return -22;
    P.source = sources[0];
  } else if (P.kind == ORC_PPR) {
    P.teleport.assign(sources, sources + n);
    std::sort(P.teleport.begin(), P.teleport.end());
    P.teleport.erase(std::unique(P.teleport.begin(), P.teleport.end()), P.teleport.end());
  } else {
    return -22;
  }
  init_program(g, P);
  return 0;
}

int orc_last_reexec(const orc_graph* g, int prog) {
  if (prog < 0 || prog >= (int)g->progs.size()) return -22;
  return g->progs[prog].last_reexec;
}

uint64_t orc_live_edges(const orc_graph* g, uint32_t* src, uint32_t* dst, uint32_t* w) {
  uint64_t n = 0;
  for (uint64_t t = 0; t < g->tiles.size(); ++t) {
    const Tile& tl = g->tiles[t];
    const uint64_t xb = (t % g->T) * (uint64_t)g->vpt, yb = (t / g->T) * (uint64_t)g->vpt;
    for (size_t i = 0; i < tl.s.size(); ++i) {
      if (tl.del[i]) continue;
      if (src) {
        src[n] = (uint32_t)(xb + tl.s[i]);
        dst[n] = (uint32_t)(yb + tl.d[i]);
        if (w) w[n] = g->weighted ? tl.w[i] : 0;
      }
      ++n;
    }
  }
  return n;
}

void orc_degrees(const orc_graph* g, uint32_t* outdeg, uint32_t* indeg) {
  std::memcpy(outdeg, g->outdeg.data(), g->V * 4);
  std::memcpy(indeg, g->indeg.data(), g->V * 4);
}

// ----------------------------------------------------------- checkers
namespace {
struct CSR {
  std::vector<uint64_t> off;
  std::vector<uint32_t> adj;
  std::vector<uint32_t> wt;
};
CSR build_csr(uint64_t V, uint64_t E, const uint32_t* a, const uint32_t* b, const uint32_t* w) {
  CSR c;
  c.off.assign(V + 1, 0);
  for (uint64_t i = 0; i < E; ++i) c.off[a[i] + 1]++;
  for (uint64_t v = 0; v < V; ++v) c.off[v + 1] += c.off[v];
  c.adj.resize(E);
  if (w) c.wt.resize(E);
  std::vector<uint64_t> pos(c.off.begin(), c.off.end() - 1);
  for (uint64_t i = 0; i < E; ++i) {
    uint64_t p = pos[a[i]]++;
    c.adj[p] = b[i];
    if (w) c.wt[p] = w[i];
  }
  return c;
}

// fn(i0, i1) on `threads` contiguous pieces of [0, n), one std::thread each.
void par_ranges(int threads, uint64_t n, const std::function<void(uint64_t, uint64_t)>& fn) {
  if (threads <= 1 || n < (1u << 16)) {
    fn(0, n);
    return;
  }
  std::vector<std::thread> ts;
  for (int t = 0; t < threads; ++t) {
    const uint64_t a = n * (uint64_t)t / (uint64_t)threads;
    const uint64_t b = n * (uint64_t)(t + 1) / (uint64_t)threads;
    ts.emplace_back([&fn, a, b] { fn(a, b); });
  }
  for (auto& t : ts) t.join();
}

// The same rows built by several threads: counts and cursors by relaxed
// atomics, so the order inside a row depends on the schedule (the checkers
// that use it do not depend on that order).  One thread = build_csr.
CSR build_csr_mt(uint64_t V, uint64_t E, const uint32_t* a, const uint32_t* b,
                 const uint32_t* w, int threads) {
  if (threads <= 1) return build_csr(V, E, a, b, w);
  CSR c;
  std::vector<uint64_t> cur(V, 0);
  par_ranges(threads, E, [&](uint64_t i0, uint64_t i1) {
    for (uint64_t i = i0; i < i1; ++i) __atomic_fetch_add(&cur[a[i]], 1ull, __ATOMIC_RELAXED);
  });
  c.off.assign(V + 1, 0);
  for (uint64_t v = 0; v < V; ++v) {
    c.off[v + 1] = c.off[v] + cur[v];
    cur[v] = c.off[v];
  }
  c.adj.resize(E);
  if (w) c.wt.resize(E);
  par_ranges(threads, E, [&](uint64_t i0, uint64_t i1) {
    for (uint64_t i = i0; i < i1; ++i) {
      const uint64_t p = __atomic_fetch_add(&cur[a[i]], 1ull, __ATOMIC_RELAXED);
      c.adj[p] = b[i];
      if (w) c.wt[p] = w[i];
    }
  });
  return c;
}

std::vector<uint32_t> out_degrees_mt(uint64_t V, uint64_t E, const uint32_t* src, int threads) {
  std::vector<uint32_t> deg(V, 0);
  par_ranges(threads, E, [&](uint64_t i0, uint64_t i1) {
    for (uint64_t i = i0; i < i1; ++i) __atomic_fetch_add(&deg[src[i]], 1u, __ATOMIC_RELAXED);
  });
  return deg;
}
}  // namespace

namespace {
int g_checker_threads = 1;  // threads that build the checkers' rows
}

void orc_checker_threads(int n) { g_checker_threads = n < 1 ? 1 : n; }

void orc_check_bfs(uint64_t V, uint64_t E, const uint32_t* src, const uint32_t* dst,
                   uint32_t source, uint32_t* out) {
  CSR c = build_csr_mt(V, E, src, dst, nullptr, g_checker_threads);
  std::fill(out, out + V, INF);
  std::vector<uint32_t> q;
  q.reserve(V);
  out[source] = 0;
  q.push_back(source);
  for (size_t h = 0; h < q.size(); ++h) {
    uint32_t u = q[h];
    for (uint64_t p = c.off[u]; p < c.off[u + 1]; ++p) {
      uint32_t v = c.adj[p];
      if (out[v] == INF) {
        out[v] = out[u] + 1;
        q.push_back(v);
      }
    }
  }
}

void orc_check_sssp(uint64_t V, uint64_t E, const uint32_t* src, const uint32_t* dst,
                    const uint32_t* w, uint32_t source, uint32_t* out) {
  CSR c = build_csr_mt(V, E, src, dst, w, g_checker_threads);
  std::vector<uint64_t> dist(V, UINT64_MAX);
  using Item = std::pair<uint64_t, uint32_t>;
  std::priority_queue<Item, std::vector<Item>, std::greater<Item>> pq;
  dist[source] = 0;
  pq.push({0, source});
  while (!pq.empty()) {
    auto [d, u] = pq.top();
    pq.pop();
    if (d != dist[u]) continue;
    for (uint64_t p = c.off[u]; p < c.off[u + 1]; ++p) {
      uint32_t v = c.adj[p];
      uint64_t nd = d + c.wt[p];
      if (nd < dist[v]) {
        dist[v] = nd;
        pq.push({nd, v});
      }
    }
  }
  for (uint64_t v = 0; v < V; ++v) out[v] = dist[v] >= INF ? INF : (uint32_t)dist[v];
}

// Queue BFS / Dijkstra from several sources over one set of rows (built
// once); the sources are split over `threads` threads, out is n x V.
namespace {
void bfs_rows(const CSR& c, uint64_t V, uint32_t source, uint32_t* out) {
  std::fill(out, out + V, INF);
  std::vector<uint32_t> q;
  q.reserve(1024);
  out[source] = 0;
  q.push_back(source);
  for (size_t h = 0; h < q.size(); ++h) {
    const uint32_t u = q[h];
    for (uint64_t p = c.off[u]; p < c.off[u + 1]; ++p) {
      const uint32_t v = c.adj[p];
      if (out[v] == INF) {
        out[v] = out[u] + 1;
        q.push_back(v);
      }
    }
  }
}

void dijkstra_rows(const CSR& c, uint64_t V, uint32_t source, uint32_t* out) {
  std::vector<uint64_t> dist(V, UINT64_MAX);
  using Item = std::pair<uint64_t, uint32_t>;
  std::priority_queue<Item, std::vector<Item>, std::greater<Item>> pq;
  dist[source] = 0;
  pq.push({0, source});
  while (!pq.empty()) {
    auto [d, u] = pq.top();
    pq.pop();
    if (d != dist[u]) continue;
    for (uint64_t p = c.off[u]; p < c.off[u + 1]; ++p) {
      const uint32_t v = c.adj[p];
      const uint64_t nd = d + c.wt[p];
      if (nd < dist[v]) {
        dist[v] = nd;
        pq.push({nd, v});
      }
    }
  }
  for (uint64_t v = 0; v < V; ++v) out[v] = dist[v] >= INF ? INF : (uint32_t)dist[v];
}
}  // namespace

void orc_check_multi(uint64_t V, uint64_t E, const uint32_t* src, const uint32_t* dst,
                     const uint32_t* w, const uint32_t* sources, uint32_t n, int threads,
                     uint32_t* out) {
  CSR c = build_csr_mt(V, E, src, dst, w, threads);
  std::atomic<uint32_t> next{0};
  const int nt = std::max(1, std::min<int>(threads, (int)n));
  std::vector<std::thread> ts;
  for (int t = 0; t < nt; ++t)
    ts.emplace_back([&] {
      for (uint32_t k = next++; k < n; k = next++) {
        if (w) dijkstra_rows(c, V, sources[k], out + (uint64_t)k * V);
        else bfs_rows(c, V, sources[k], out + (uint64_t)k * V);
      }
    });
  for (auto& t : ts) t.join();
}

// label[v] = min({v} U {u : u reaches v}) over directed paths.
void orc_check_cc(uint64_t V, uint64_t E, const uint32_t* src, const uint32_t* dst,
                  uint32_t* out) {
  orc_check_cc_mt(V, E, src, dst, g_checker_threads, out);
}

// The same labels; the rows are built by `threads` threads, the search is
// one depth-first sweep in increasing root order (each vertex is labelled by
// the first, i.e. smallest, root that reaches it).
void orc_check_cc_mt(uint64_t V, uint64_t E, const uint32_t* src, const uint32_t* dst,
                     int threads, uint32_t* out) {
  CSR c = build_csr_mt(V, E, src, dst, nullptr, threads);
  std::fill(out, out + V, INF);
  std::vector<uint32_t> stack;
  for (uint64_t u = 0; u < V; ++u) {
    if (out[u] != INF) continue;
    out[u] = (uint32_t)u;
    stack.push_back((uint32_t)u);
    while (!stack.empty()) {
      uint32_t a = stack.back();
      stack.pop_back();
      for (uint64_t p = c.off[a]; p < c.off[a + 1]; ++p) {
        uint32_t v = c.adj[p];
        if (out[v] == INF) {
          out[v] = (uint32_t)u;
          stack.push_back(v);
        }
      }
    }
  }
}

uint32_t orc_prd_f64(uint64_t V, uint64_t E, const uint32_t* src, const uint32_t* dst,
                     uint32_t max_iterations, double delta, double* rank) {
  return orc_prd_f64_mt(V, E, src, dst, max_iterations, delta, 1, rank);
}

// pagerank-delta.cc:127-174 in FP64 from init (rank 0, delta 1, all
// active).  Each vertex's sum is taken over its in-edges in row order by
// one thread, so the threads only split the vertex range.
uint32_t orc_prd_f64_mt(uint64_t V, uint64_t E, const uint32_t* src, const uint32_t* dst,
                        uint32_t max_iterations, double delta, int threads, double* rank) {
  CSR in = build_csr_mt(V, E, dst, src, nullptr, threads);  // in-edges
  std::vector<uint32_t> outdeg = out_degrees_mt(V, E, src, threads);
  std::vector<double> d(V, 1.0), s(V, 0.0);
  std::vector<uint8_t> act(V, 1), nact(V, 0);
  std::fill(rank, rank + V, 0.0);
  uint32_t it = 0;
  bool first = true;
  while (it < max_iterations) {
    par_ranges(threads, V, [&](uint64_t v0, uint64_t v1) {
      for (uint64_t v = v0; v < v1; ++v) {
        double acc = 0.0;
        for (uint64_t p = in.off[v]; p < in.off[v + 1]; ++p) {
          uint32_t u = in.adj[p];
          if (act[u]) acc += d[u] / (double)outdeg[u];
        }
        s[v] = acc;
      }
    });
    std::atomic<bool> conv{true};
    par_ranges(threads, V, [&](uint64_t v0, uint64_t v1) {
      bool c = true;
      for (uint64_t v = v0; v < v1; ++v) {
        nact[v] = 0;
        double dv;
        if (first) {
          dv = (1. - ALPHA) + ALPHA * s[v];
          rank[v] += dv;
          dv -= 1.0;
        } else {
          if (s[v] == 0.0) continue;
          dv = ALPHA * s[v];
          rank[v] += dv;
        }
        d[v] = dv;
        if (std::abs(dv) > delta * rank[v]) {
          nact[v] = 1;
          c = false;
        }
      }
      if (!c) conv = false;
    });
    ++it;
    first = false;
    act.swap(nact);
    if (conv) break;
  }
  return it;
}

// One FP64 sweep of the PageRank fixed point over FP32 ranks:
// out[0] = sum_v |r_v - (1-alpha) - alpha * sum_{u->v} r_u / outdeg(u)|,
// out[1] = sum_v |r_v|.  A PageRank-delta run that pushes every delta it
// computes ends on that fixed point (pagerank-delta.cc:142-174: rank_v =
// (1-alpha) + alpha * sum_u (1 + pushed deltas of u) / outdeg(u)); deltas it
// drops below the threshold leave a residual of at most alpha * delta *
// sum_u r_u, plus the FP32 rounding of the ranks.
void orc_prd_residual(uint64_t V, uint64_t E, const uint32_t* src, const uint32_t* dst,
                      const float* rank, double alpha, int threads, double* out) {
  CSR in = build_csr_mt(V, E, dst, src, nullptr, threads);
  std::vector<uint32_t> outdeg = out_degrees_mt(V, E, src, threads);
  const int nt = threads < 1 ? 1 : threads;
  std::vector<double> part(2 * (size_t)nt, 0.0);
  std::atomic<int> slot{0};
  par_ranges(nt, V, [&](uint64_t v0, uint64_t v1) {
    double res = 0.0, tot = 0.0;
    for (uint64_t v = v0; v < v1; ++v) {
      double acc = 0.0;
      for (uint64_t p = in.off[v]; p < in.off[v + 1]; ++p) {
        const uint32_t u = in.adj[p];
        acc += (double)rank[u] / (double)outdeg[u];
      }
      res += std::abs((double)rank[v] - (1.0 - alpha) - alpha * acc);
      tot += std::abs((double)rank[v]);
    }
    const int k = slot.fetch_add(1);
    part[2 * k] = res;
    part[2 * k + 1] = tot;
  });
  out[0] = out[1] = 0.0;
  for (int k = 0; k < slot.load(); ++k) {
    out[0] += part[2 * k];
    out[1] += part[2 * k + 1];
  }
}

uint32_t orc_ppr_f64(uint64_t V, uint64_t E, const uint32_t* src, const uint32_t* dst,
                     const uint32_t* teleport, uint32_t n_teleport, double alpha, double tol,
                     uint32_t max_iterations, double* rank) {
  CSR in = build_csr(V, E, dst, src, nullptr);
  std::vector<uint32_t> outdeg(V, 0);
  for (uint64_t i = 0; i < E; ++i) outdeg[src[i]]++;
  std::vector<uint32_t> set(teleport, teleport + n_teleport);
  std::sort(set.begin(), set.end());
  set.erase(std::unique(set.begin(), set.end()), set.end());
  std::vector<double> t(V, 0.0), r(V, 0.0), nr(V);
  for (uint32_t v : set) t[v] = 1.0 / (double)set.size();
  uint32_t it = 0;
  while (it < max_iterations) {
    double diff = 0.0;
    for (uint64_t v = 0; v < V; ++v) {
      double acc = 0.0;
      for (uint64_t p = in.off[v]; p < in.off[v + 1]; ++p) {
        const uint32_t u = in.adj[p];
        acc += r[u] / (double)outdeg[u];
      }
      nr[v] = (1.0 - alpha) * t[v] + alpha * acc;
      diff = std::max(diff, std::abs(nr[v] - r[v]));
    }
    r.swap(nr);
    ++it;
    if (diff < tol) break;
  }
  std::copy(r.begin(), r.end(), rank);
  return it;
}

uint32_t orc_pagerank_f64(uint64_t V, uint64_t E, const uint32_t* src, const uint32_t* dst,
                          double alpha, double tol, uint32_t max_iterations, double* rank) {
  CSR in = build_csr(V, E, dst, src, nullptr);
  std::vector<uint32_t> outdeg(V, 0);
  for (uint64_t i = 0; i < E; ++i) outdeg[src[i]]++;
  std::vector<double> r(V, 1.0), nr(V);
  uint32_t it = 0;
  while (it < max_iterations) {
    double diff = 0.0;
    for (uint64_t v = 0; v < V; ++v) {
      double acc = 0.0;
      for (uint64_t p = in.off[v]; p < in.off[v + 1]; ++p) {
        uint32_t u = in.adj[p];
        acc += r[u] / (double)outdeg[u];
      }
      nr[v] = (1.0 - alpha) + alpha * acc;
      diff = std::max(diff, std::abs(nr[v] - r[v]));
    }
    r.swap(nr);
    ++it;
    if (diff < tol) break;
  }
  std::copy(r.begin(), r.end(), rank);
  return it;
}

}  // extern "C"
