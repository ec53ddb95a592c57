// tagore_oracle.cpp -- CPU restatement of Tagore's kNN-graph construction (GNN-Descent)
// and CFS (Collect-Filter-Store) pruning rules.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by tests/, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the product
// library (tagore_amd/), which must fail loudly if its HIP build is missing.
//
// Every function names the reference file:line it restates (paths relative to the
// Tagore source tree).  Where the reference is racy or reads uninitialised memory the
// deterministic decision taken is stated next to the code; the GPU kernels in
// tagore_amd/csrc implement the same decisions, so pruning output is compared
// bit-for-bit and the exact-mode kNN graph is compared bit-for-bit too.
//
// Conventions
//   * dist(a,b) is the canonical FP32 squared L2 (cdist below); every distance that reaches a
//     list, a sort or a rule is cdist.
//   * "sort" means ascending on (key, id); keys are non-negative floats (or u32 detours).
//   * INV = 0xFFFFFFFF marks an invalid id.
//   * Built with -ffp-contract=off so the rule arithmetic is evaluated exactly as written.
//
// Parity status (what pins each part) is recorded in the header of each section.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <vector>
#include <omp.h>

namespace {

constexpr uint32_t INV = 0xFFFFFFFFu;
constexpr float FINF = std::numeric_limits<float>::infinity();
constexpr int RING = 480;          // SELECT_CAND, Tagore_src.cuh:35
constexpr int PATH_EXPANSIONS = 10;  // Tagore_src.cu:1261
constexpr int EP_EXPANSIONS = 30;    // Tagore_src.cu:1961
constexpr int CENTROID_CHUNK = 1024;

// ---------------------------------------------------------------------------------------
// Canonical FP32 squared L2.  64 partial sums p[l] over i == l (mod 64), each folded in
// ascending i with fmaf(t, t, acc); then the four partials of each run l = 4t..4t+3 are
// summed as q[t] = (p[4t] + p[4t+1]) + (p[4t+2] + p[4t+3]), and the tree q[t] += q[t+s]
// for s = 8..1 gives the result (six levels in all).
// Pinned by: closed form vs FP64 (tests/test_oracle.py), symmetry, D=96/128/960.
// (Replaces the FP16 __hfma2 loops, e.g. Tagore_src.cu:261-286, 1429-1446.)
// ---------------------------------------------------------------------------------------
inline float cdist(const float* a, const float* b, int D) {
  float p[64];
  for (int l = 0; l < 64; ++l) p[l] = 0.0f;
  for (int i = 0; i < D; ++i) {
    float t = a[i] - b[i];
    p[i & 63] = std::fmaf(t, t, p[i & 63]);
  }
  float q[16];
  for (int t = 0; t < 16; ++t) q[t] = (p[4 * t] + p[4 * t + 1]) + (p[4 * t + 2] + p[4 * t + 3]);
  for (int s = 8; s >= 1; s >>= 1)
    for (int l = 0; l < s; ++l) q[l] = q[l] + q[l + s];
  return q[0];
}

struct Ent {
  float d;
  uint32_t id;
};
inline bool ent_lt(const Ent& a, const Ent& b) {
  return a.d < b.d || (a.d == b.d && a.id < b.id);
}
inline bool ent_eq(const Ent& a, const Ent& b) { return a.d == b.d && a.id == b.id; }

inline uint64_t ent_key(float d, uint32_t id) {
  uint32_t bits;
  std::memcpy(&bits, &d, 4);
  return (uint64_t(bits) << 32) | id;
}
inline float key_d(uint64_t k) {
  uint32_t bits = uint32_t(k >> 32);
  float d;
  std::memcpy(&d, &bits, 4);
  return d;
}
inline uint32_t key_id(uint64_t k) { return uint32_t(k); }

// Published SplitMix64 (Steele, Lea, Flood, OOPSLA 2014).
inline uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
inline uint64_t stream_seed(uint64_t seed, uint64_t idx) {
  uint64_t s = seed ^ (idx * 0xD6E8FEB86659FD93ULL);
  return splitmix64(s);
}
inline uint32_t bounded(uint64_t x, uint32_t n) {
  return uint32_t(((x >> 32) * uint64_t(n)) >> 32);
}

// Rules of the Filter step (Tagore_src.cu:1488-1495); returns < 0 to prune.
//   res = d(p,c), dis = d(u,c), cur = d(u,p).
inline float rule_dist(float res, float dis, float cur, float th) {
  (void)cur;
  return th * res - dis;
}
inline float rule_angle(float res, float dis, float cur, float th) {
  return th - ((dis + cur) - res) / 2.0f / std::sqrt(dis * cur);
}
inline float apply_rule(int rule, float res, float dis, float cur, float th) {
  return rule == 0 ? rule_dist(res, dis, cur, th) : rule_angle(res, dis, cur, th);
}

// Per-destination push lists: every push (dest, key); the reduction keeps, per dest, the
// `cap` smallest keys (multiset order), which replaces the reference's atomicAdd slot
// order plus "replace if closer" (Tagore_src.cu:380-390, 1465-1475, 1697-1707).
struct PushList {
  std::vector<std::vector<std::pair<uint32_t, uint64_t>>> per_thread;
  explicit PushList(int nthreads) : per_thread(nthreads) {}
  void push(int tid, uint32_t dest, uint64_t key) { per_thread[tid].push_back({dest, key}); }
  // Returns for every dest the sorted list of its smallest `cap` keys (cap==0: all).
  void reduce(uint32_t N, uint32_t cap, std::vector<uint32_t>& off,
              std::vector<uint64_t>& keys) {
    off.assign(N + 1, 0);
    for (auto& v : per_thread)
      for (auto& e : v) off[e.first + 1]++;
    for (uint32_t i = 0; i < N; ++i) off[i + 1] += off[i];
    std::vector<uint64_t> all(off[N]);
    std::vector<uint32_t> cur(off.begin(), off.end() - 1);
    for (auto& v : per_thread)
      for (auto& e : v) all[cur[e.first]++] = e.second;
    for (auto& v : per_thread) std::vector<std::pair<uint32_t, uint64_t>>().swap(v);
    std::vector<uint32_t> noff(N + 1, 0);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t v = 0; v < int64_t(N); ++v) {
      std::sort(all.begin() + off[v], all.begin() + off[v + 1]);
      uint32_t n = off[v + 1] - off[v];
      noff[v + 1] = cap ? std::min(n, cap) : n;
    }
    for (uint32_t i = 0; i < N; ++i) noff[i + 1] += noff[i];
    keys.assign(noff[N], 0);
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < int64_t(N); ++v) {
      uint32_t n = noff[v + 1] - noff[v];
      std::copy(all.begin() + off[v], all.begin() + off[v] + n, keys.begin() + noff[v]);
    }
    off.swap(noff);
  }
};

// ---------------------------------------------------------------------------------------
// Beam search shared by cal_ep_gpu (Tagore_src.cu:1889-2058) and collect_path
// (Tagore_src.cu:1209-1349).  Spec (decisions A.0 #9, #12):
//   S = G[start] with distances (on_eval is called for each, in G order), sorted;
//   B = S[0..min(TOPM,K)) padded with (+inf, INV, explored);  dedup set = ids(B) u ids(S);
//   repeat <= max_exp: x = first entry of B that is unexplored and finite (stop if none);
//     new_j = (+inf, y) if y = G[x][j] is in the dedup set, else (dist(y), y) and on_eval;
//     W = sort(B u new) by (d,id); dedup set = ids(W); B = W[0..TOPM).
// ---------------------------------------------------------------------------------------
template <class DistFn, class OnEval>
void beam_search(const uint32_t* G, int K, int TOPM, uint32_t start, int max_exp,
                 DistFn&& dfn, OnEval&& on_eval, std::vector<Ent>& B,
                 std::vector<uint8_t>& explored, uint64_t* n_eval, uint64_t* n_adj) {
  std::vector<Ent> S(K);
  for (int j = 0; j < K; ++j) {
    uint32_t y = G[uint64_t(start) * K + j];
    S[j] = {dfn(y), y};
    on_eval(S[j]);
  }
  if (n_eval) *n_eval += K;
  if (n_adj) *n_adj += 1;
  std::sort(S.begin(), S.end(), ent_lt);
  const int nb = std::min(TOPM, K);
  B.assign(TOPM, Ent{FINF, INV});
  explored.assign(TOPM, 1);
  for (int i = 0; i < nb; ++i) {
    B[i] = S[i];
    explored[i] = 0;
  }
  std::vector<uint32_t> dset;
  dset.reserve(TOPM + K);
  for (auto& e : S) dset.push_back(e.id);
  if (nb < TOPM) dset.push_back(INV);
  std::sort(dset.begin(), dset.end());

  std::vector<Ent> W(TOPM + K);
  std::vector<uint8_t> Wexp(TOPM + K);
  std::vector<int> idx(TOPM + K);
  for (int r = 0; r < max_exp; ++r) {
    int xi = -1;
    for (int i = 0; i < TOPM; ++i)
      if (!explored[i] && B[i].d < FINF) {
        xi = i;
        break;
      }
    if (xi < 0) break;
    explored[xi] = 1;
    const uint32_t x = B[xi].id;
    if (n_adj) *n_adj += 1;
    for (int i = 0; i < TOPM; ++i) {
      W[i] = B[i];
      Wexp[i] = explored[i];
    }
    for (int j = 0; j < K; ++j) {
      uint32_t y = G[uint64_t(x) * K + j];
      Ent e{FINF, y};
      if (!std::binary_search(dset.begin(), dset.end(), y)) {
        e.d = dfn(y);
        on_eval(e);
        if (n_eval) *n_eval += 1;
      }
      W[TOPM + j] = e;
      Wexp[TOPM + j] = 0;
    }
    for (int i = 0; i < TOPM + K; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return ent_lt(W[a], W[b]); });
    dset.clear();
    for (int i = 0; i < TOPM + K; ++i) dset.push_back(W[i].id);
    std::sort(dset.begin(), dset.end());
    for (int i = 0; i < TOPM; ++i) {
      B[i] = W[idx[i]];
      explored[i] = Wexp[idx[i]];
    }
  }
}

// Sort a pool by (d,id) and invalidate adjacent repeated ids (compared against the
// sorted array before any invalidation; Tagore_src.cu:1345-1348, 1391-1394).
void sort_dedup_pool(std::vector<Ent>& P) {
  std::sort(P.begin(), P.end(), ent_lt);
  std::vector<uint32_t> ids(P.size());
  for (size_t i = 0; i < P.size(); ++i) ids[i] = P[i].id;
  for (size_t i = 1; i < P.size(); ++i)
    if (ids[i] == ids[i - 1]) P[i].id = INV;
}

// ---------------------------------------------------------------------------------------
// Filter (Filter_path, Tagore_src.cu:1399-1485; also the filter loop of filter_reverse,
// :1561-1626).  Spec A.4.  Returns the kept (id, d) list: all selections except the last
// one appended (FD reached: FD-1 kept; selection from index L-1: that one dropped;
// pool exhausted: every real selection kept).  No valid first candidate: degree 0.
// n_eval counts the rule evaluations (one distance each) of this incremental model.
// ---------------------------------------------------------------------------------------
void filter_pool(std::vector<Ent>& P, uint32_t u, int FD, int rule, float th, const float* X,
                 int D, std::vector<Ent>& out, uint64_t* n_eval) {
  out.clear();
  const int L = int(P.size());
  int t = 0;
  while (t < L && (P[t].id == u || P[t].id == INV)) ++t;
  if (t >= L) return;
  std::vector<Ent> sel;
  sel.push_back(P[t]);
  float cur = P[t].d;
  ++t;
  bool garbage_last = false;
  while (int(sel.size()) < FD) {
    const Ent p = sel.back();
    const float* xp = X + uint64_t(p.id) * D;
    for (int i = t; i < L; ++i) {
      if (P[i].id == INV) continue;
      if (P[i].id == p.id) {
        P[i].id = INV;
        continue;
      }
      float res = cdist(xp, X + uint64_t(P[i].id) * D, D);
      if (n_eval) *n_eval += 1;
      if (apply_rule(rule, res, P[i].d, cur, th) < 0.0f) P[i].id = INV;
    }
    while (t < L && P[t].id == INV) ++t;
    if (t < L) {
      sel.push_back(P[t]);
      cur = P[t].d;
    } else {
      garbage_last = true;
      sel.push_back(Ent{FINF, INV});
    }
    ++t;
    if (t >= L) break;
  }
  (void)garbage_last;
  // Output drops the last appended entry (Tagore_src.cu:1462-1463, 1477).
  if (sel.size() >= 1) sel.pop_back();
  out = sel;
}

struct PruneCounters {
  uint64_t collect_evals = 0, filter_evals = 0, reverse_evals = 0, adj_rows = 0;
  uint64_t ep_evals = 0;
};

// collect_2hop (Tagore_src.cu:1352-1396; spec A.6): appends the first 480 two-hop ids,
// slot i holding G[G[u][i / K]][i % K] (rows of the first neighbours, whole rows in list
// order), each with its canonical distance to u.
template <class DistFn>
void collect_2hop_pool(const uint32_t* G, int K, int64_t u, DistFn&& dfn, std::vector<Ent>& P) {
  for (int i = 0; i < RING; ++i) {
    uint32_t v = G[u * K + i / K];
    uint32_t y = G[uint64_t(v) * K + (i % K)];
    P.push_back({dfn(y), y});
  }
}

// Reverse phase (filter_reverse, Tagore_src.cu:1537-1657; spec A.7).  For every node u:
// C = its forward row (fdeg[u] entries, in order) followed by its reverse list, the K
// smallest (d, src) pushes addressed to u in ascending order (cap K: the reverse_num cap of
// the Store step, Tagore_src.cu:1465-1475).  If |C| > FD, C is sorted, deduplicated and
// filtered again with the reverse rule (at most FD-1 kept, :1561-1626); otherwise the row is
// the forward row followed by the reverse ids not already in it, in list order
// (:1628-1649).  Consumes `pushes`; returns the filter's rule evaluations.
uint64_t reverse_phase(const float* X, int64_t N, int D, int K, int FD, int rev_rule, float th,
                       const uint32_t* fdeg, const uint32_t* fids, const float* fdd,
                       PushList& pushes, uint32_t* out_deg, uint32_t* out_ids,
                       uint32_t* rev_cnt, uint64_t* rev_keys) {
  const int nthreads = omp_get_max_threads();
  std::vector<uint64_t> cnt_f(nthreads, 0);
  std::vector<uint32_t> roff;
  std::vector<uint64_t> rkeys;
  pushes.reduce(uint32_t(N), uint32_t(K), roff, rkeys);
#pragma omp parallel
  {
    const int tid = omp_get_thread_num();
    std::vector<Ent> C, out;
#pragma omp for schedule(dynamic, 64)
    for (int64_t u = 0; u < N; ++u) {
      const uint32_t nr = roff[u + 1] - roff[u];
      const uint32_t deg = fdeg[u];
      C.clear();
      for (uint32_t k = 0; k < deg; ++k) C.push_back({fdd[u * FD + k], fids[u * FD + k]});
      for (uint32_t k = 0; k < nr; ++k) {
        uint64_t key = rkeys[roff[u] + k];
        C.push_back({key_d(key), key_id(key)});
      }
      uint32_t* row = out_ids + u * FD;
      for (int k = 0; k < FD; ++k) row[k] = INV;
      if (int(C.size()) > FD) {
        sort_dedup_pool(C);
        uint64_t nf = 0;
        filter_pool(C, uint32_t(u), FD, rev_rule, th, X, D, out, &nf);
        cnt_f[tid] += nf;
        out_deg[u] = uint32_t(out.size());
        for (size_t k = 0; k < out.size(); ++k) row[k] = out[k].id;
      } else {
        uint32_t n = 0;
        for (uint32_t k = 0; k < deg; ++k) row[n++] = fids[u * FD + k];
        for (uint32_t k = 0; k < nr; ++k) {
          uint32_t w = key_id(rkeys[roff[u] + k]);
          bool found = false;
          for (uint32_t j = 0; j < deg; ++j)
            if (fids[u * FD + j] == w) {
              found = true;
              break;
            }
          if (!found) row[n++] = w;
        }
        out_deg[u] = n;
      }
    }
  }
  if (rev_cnt || rev_keys) {
    for (int64_t u = 0; u < N; ++u) {
      uint32_t nr = roff[u + 1] - roff[u];
      if (rev_cnt) rev_cnt[u] = nr;
      if (rev_keys)
        for (uint32_t k = 0; k < nr; ++k) rev_keys[u * K + k] = rkeys[roff[u] + k];
    }
  }
  uint64_t total = 0;
  for (int t = 0; t < nthreads; ++t) total += cnt_f[t];
  return total;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------------------
// Basic arithmetic
// ---------------------------------------------------------------------------------------
float or_dist(const float* a, const float* b, int D) { return cdist(a, b, D); }

void or_dist_pairs(const float* X, int D, const uint32_t* a, const uint32_t* b, int64_t n,
                   float* out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i)
    out[i] = cdist(X + uint64_t(a[i]) * D, X + uint64_t(b[i]) * D, D);
}

// Deterministic centroid: FP64 sums over chunks of 1024 rows in index order, chunk partials
// summed in order, divided by N in FP64 and rounded once to FP32.  (Reference: FP32 host
// loop pybind.cu:40-51; decision: deterministic FP64 restatement shared with the GPU.)
void or_centroid(const float* X, int64_t N, int D, float* c) {
  const int64_t nch = (N + CENTROID_CHUNK - 1) / CENTROID_CHUNK;
  std::vector<double> part(size_t(nch) * D, 0.0);
#pragma omp parallel for schedule(static)
  for (int64_t ch = 0; ch < nch; ++ch) {
    int64_t lo = ch * CENTROID_CHUNK, hi = std::min<int64_t>(N, lo + CENTROID_CHUNK);
    for (int j = 0; j < D; ++j) {
      double s = 0.0;
      for (int64_t i = lo; i < hi; ++i) s += double(X[i * D + j]);
      part[size_t(ch) * D + j] = s;
    }
  }
  for (int j = 0; j < D; ++j) {
    double s = 0.0;
    for (int64_t ch = 0; ch < nch; ++ch) s += part[size_t(ch) * D + j];
    c[j] = float(s / double(N));
  }
}

// Exact kNN in canonical FP32 (rows sorted by (d,id), self excluded).  Used to build
// pruning inputs; pinned against numpy FP64 on tiny inputs.
int or_knn_bruteforce(const float* X, int64_t N, int D, int K, uint32_t* ids, float* dists) {
  if (K >= N) return 1;
#pragma omp parallel
  {
    std::vector<Ent> buf(N);
#pragma omp for schedule(dynamic, 16)
    for (int64_t u = 0; u < N; ++u) {
      int64_t m = 0;
      for (int64_t v = 0; v < N; ++v)
        if (v != u) buf[m++] = {cdist(X + u * D, X + v * D, D), uint32_t(v)};
      std::partial_sort(buf.begin(), buf.begin() + K, buf.begin() + m, ent_lt);
      for (int j = 0; j < K; ++j) {
        ids[u * K + j] = buf[j].id;
        dists[u * K + j] = buf[j].d;
      }
    }
  }
  return 0;
}

// FP64 exact neighbours of the query rows (recall judge).  Output: k ids sorted by
// (d64,id) and the k-th FP64 distance (for the tie rule).
int or_knn_f64(const float* X, int64_t N, int D, const uint32_t* q, int64_t nq, int k,
               uint32_t* ids, double* kth) {
  if (k >= N) return 1;
#pragma omp parallel
  {
    std::vector<std::pair<double, uint32_t>> buf(N);
#pragma omp for schedule(dynamic, 4)
    for (int64_t qi = 0; qi < nq; ++qi) {
      const uint32_t u = q[qi];
      int64_t m = 0;
      for (int64_t v = 0; v < N; ++v) {
        if (v == u) continue;
        double s = 0.0;
        for (int j = 0; j < D; ++j) {
          double t = double(X[uint64_t(u) * D + j]) - double(X[v * D + j]);
          s += t * t;
        }
        buf[m++] = {s, uint32_t(v)};
      }
      std::partial_sort(buf.begin(), buf.begin() + k, buf.begin() + m);
      for (int j = 0; j < k; ++j) ids[qi * k + j] = buf[j].second;
      kth[qi] = buf[k - 1].first;
    }
  }
  return 0;
}

// recall@k with the tie rule: a returned id is a hit when its FP64 distance is <= the k-th
// exact FP64 distance.  G rows have stride K; the first k entries are judged.
double or_recall(const float* X, int64_t N, int D, const uint32_t* G, int K,
                 const uint32_t* q, int64_t nq, int k, const double* kth) {
  (void)N;
  double tot = 0.0;
#pragma omp parallel for reduction(+ : tot) schedule(static)
  for (int64_t qi = 0; qi < nq; ++qi) {
    const uint32_t u = q[qi];
    int hits = 0;
    for (int j = 0; j < k; ++j) {
      uint32_t v = G[uint64_t(u) * K + j];
      if (v == INV || v == u) continue;
      double s = 0.0;
      for (int t = 0; t < D; ++t) {
        double z = double(X[uint64_t(u) * D + t]) - double(X[uint64_t(v) * D + t]);
        s += z * z;
      }
      if (s <= kth[qi]) ++hits;
    }
    tot += double(std::min(hits, k)) / k;
  }
  return tot / double(nq);
}

// ---------------------------------------------------------------------------------------
// Entry point (cal_ep_gpu, Tagore_src.cu:1889-2058; spec A.2): beam search of width TOPM
// from node 0's list toward the centroid, at most 30 expansions; ep = best beam entry.
// ---------------------------------------------------------------------------------------
uint32_t or_cal_ep(const float* X, int64_t N, int D, const uint32_t* G, int K, int TOPM,
                   const float* centroid, uint64_t* n_eval) {
  (void)N;
  std::vector<Ent> B;
  std::vector<uint8_t> ex;
  auto dfn = [&](uint32_t y) { return cdist(centroid, X + uint64_t(y) * D, D); };
  beam_search(G, K, TOPM, 0, EP_EXPANSIONS, dfn, [](const Ent&) {}, B, ex, n_eval, nullptr);
  return B[0].id;
}

// ---------------------------------------------------------------------------------------
// Pruning.  index_kind: 0 path collect (NSG / Vamana / CFS 'path'), 1 two-hop collect
// (NSSG / CFS '2-hop'), 2 CAGRA one-hop rank, 3 DPG one-hop rank.
// fwd_rule (path/2-hop): 0 dist, 1 angle.  rev_rule: 0 dist, 1 angle, 2 rank.
// Inputs: X (N x D fp32), kNN graph G/GD (N x K, rows sorted by (d,id)).
// Outputs: out_deg[N], out_ids[N*FD] (row stride FD), ep (path only, else 0).
// Optional stage outputs (may be null): fwd_deg[N], fwd_ids[N*FD], fwd_d[N*FD],
// rev_cnt[N], rev_keys[N*cap] (cap = K for dist/angle, FD for rank).
// Returns 0 on success, nonzero on an invalid combination.
// ---------------------------------------------------------------------------------------
int or_prune(const float* X, int64_t N, int D, const uint32_t* G, const float* GD, int K,
             int FD, int TOPM, int index_kind, int fwd_rule, int rev_rule, float th,
             uint32_t* out_deg, uint32_t* out_ids, uint32_t* ep_out, uint32_t* fwd_deg,
             uint32_t* fwd_ids, float* fwd_d, uint32_t* rev_cnt, uint64_t* rev_keys,
             uint64_t* counters) {
  const bool rank = index_kind >= 2;
  if (rank != (rev_rule == 2)) return 2;  // A.12: layout-incompatible combination
  if (FD < 1 || K < 1 || TOPM < 1 || (rank && FD > K)) return 3;
  const int nthreads = omp_get_max_threads();
  PruneCounters pc;
  uint32_t ep = 0;
  std::vector<float> cen(D);
  if (index_kind == 0) {
    or_centroid(X, N, D, cen.data());
    ep = or_cal_ep(X, N, D, G, K, TOPM, cen.data(), &pc.ep_evals);
  }
  if (ep_out) *ep_out = ep;

  PushList pushes(nthreads);
  std::vector<uint32_t> fdeg(N);
  std::vector<uint32_t> fids(size_t(N) * FD, INV);
  std::vector<float> fdd(size_t(N) * FD, 0.0f);
  std::vector<uint64_t> cnt_c(nthreads, 0), cnt_f(nthreads, 0), cnt_a(nthreads, 0);

  if (!rank) {
    // Forward phase: Collect (path or 2-hop) + Filter + Store reverse pushes.
#pragma omp parallel
    {
      const int tid = omp_get_thread_num();
      std::vector<Ent> P, ring, out, B;
      std::vector<uint8_t> ex;
#pragma omp for schedule(dynamic, 32)
      for (int64_t u = 0; u < N; ++u) {
        const float* xu = X + u * D;
        auto dfn = [&](uint32_t y) { return cdist(xu, X + uint64_t(y) * D, D); };
        P.clear();
        for (int j = 0; j < K; ++j) P.push_back({GD[u * K + j], G[u * K + j]});
        if (index_kind == 0) {
          // collect_path (Tagore_src.cu:1209-1349; spec A.3): ring of 480 slots filled in
          // evaluation order (start list first, then expansion order and j), wrapping.
          ring.assign(RING, Ent{FINF, INV});
          uint64_t c = 0;
          auto on_eval = [&](const Ent& e) { ring[c % RING] = e; ++c; };
          uint64_t ne = 0, na = 0;
          beam_search(G, K, TOPM, ep, PATH_EXPANSIONS, dfn, on_eval, B, ex, &ne, &na);
          cnt_c[tid] += ne;
          cnt_a[tid] += na;
          for (auto& e : ring) P.push_back(e);
        } else {
          collect_2hop_pool(G, K, u, dfn, P);
          cnt_c[tid] += RING;
          cnt_a[tid] += (RING + K - 1) / K;
        }
        sort_dedup_pool(P);
        uint64_t nf = 0;
        filter_pool(P, uint32_t(u), FD, fwd_rule, th, X, D, out, &nf);
        cnt_f[tid] += nf;
        fdeg[u] = uint32_t(out.size());
        for (size_t k = 0; k < out.size(); ++k) {
          fids[u * FD + k] = out[k].id;
          fdd[u * FD + k] = out[k].d;
          pushes.push(tid, out[k].id, ent_key(out[k].d, uint32_t(u)));
        }
      }
    }
    for (int t = 0; t < nthreads; ++t) {
      pc.collect_evals += cnt_c[t];
      pc.filter_evals += cnt_f[t];
      pc.adj_rows += cnt_a[t];
    }
    pc.reverse_evals = reverse_phase(X, N, D, K, FD, rev_rule, th, fdeg.data(), fids.data(),
                                     fdd.data(), pushes, out_deg, out_ids, rev_cnt, rev_keys);
  } else {
    // One-hop rank (filter_cagra :1668-1710 / filter_dpg :1713-1785; specs A.8, A.9).
    const int d = FD;
#pragma omp parallel
    {
      const int tid = omp_get_thread_num();
      std::vector<uint32_t> det(K), pos(K), srow(K);
#pragma omp for schedule(dynamic, 64)
      for (int64_t u = 0; u < N; ++u) {
        const uint32_t* g = G + u * K;
        for (int j = 0; j < K; ++j) det[j] = 0;
        if (index_kind == 2) {
          for (int r = 0; r < K; ++r) {
            const uint32_t* gr = G + uint64_t(g[r]) * K;
            for (int j = 0; j < K; ++j) srow[j] = gr[j];
            std::sort(srow.begin(), srow.end());
            for (int j = r + 1; j < K; ++j)
              if (std::binary_search(srow.begin(), srow.end(), g[j])) det[j]++;
          }
          cnt_a[tid] += K;
        } else {
          for (int r = 0; r < K; ++r) {
            const float* xr = X + uint64_t(g[r]) * D;
            for (int j = r + 1; j < K; ++j)
              if (cdist(xr, X + uint64_t(g[j]) * D, D) < GD[u * K + j]) det[j]++;
          }
          cnt_f[tid] += uint64_t(K) * (K - 1) / 2;
        }
        for (int j = 0; j < K; ++j) pos[j] = j;
        std::sort(pos.begin(), pos.end(), [&](uint32_t a, uint32_t b) {
          return det[a] < det[b] || (det[a] == det[b] && g[a] < g[b]);
        });
        fdeg[u] = d;
        for (int i = 0; i < d; ++i) {
          fids[u * FD + i] = g[pos[i]];
          fdd[u * FD + i] = 0.0f;
          pushes.push(tid, g[pos[i]], (uint64_t(det[pos[i]]) << 32) | uint32_t(u));
        }
      }
    }
    for (int t = 0; t < nthreads; ++t) {
      pc.adj_rows += cnt_a[t];
      pc.filter_evals += cnt_f[t];
    }
    std::vector<uint32_t> roff;
    std::vector<uint64_t> rkeys;
    pushes.reduce(uint32_t(N), uint32_t(d), roff, rkeys);
    // filter_reverse_1hop (Tagore_src.cu:1812-1886; spec A.10).
#pragma omp parallel
    {
      std::vector<std::pair<uint32_t, uint32_t>> Rv;  // (id, detour)
      std::vector<uint32_t> A;
#pragma omp for schedule(dynamic, 64)
      for (int64_t u = 0; u < N; ++u) {
        const uint32_t nr = std::min<uint32_t>(roff[u + 1] - roff[u], uint32_t(d));
        const uint32_t* T = fids.data() + u * FD;
        Rv.clear();
        for (uint32_t k = 0; k < nr; ++k) {
          uint64_t key = rkeys[roff[u] + k];
          Rv.push_back({key_id(key), uint32_t(key >> 32)});
        }
        std::sort(Rv.begin(), Rv.end());
        {
          std::vector<uint32_t> orig(Rv.size());
          for (size_t i = 0; i < Rv.size(); ++i) orig[i] = Rv[i].first;
          for (size_t i = 1; i < Rv.size(); ++i)
            if (orig[i] == orig[i - 1]) Rv[i].second = INV;
        }
        std::sort(Rv.begin(), Rv.end(), [](const std::pair<uint32_t, uint32_t>& a,
                                           const std::pair<uint32_t, uint32_t>& b) {
          return a.second < b.second || (a.second == b.second && a.first < b.first);
        });
        uint32_t* row = out_ids + u * FD;
        const int h = d / 2;
        for (int i = 0; i < h; ++i) row[i] = T[i];
        const int r = int(Rv.size());
        A.clear();
        if (r < h) {
          for (auto& e : Rv) {
            bool in = false;
            for (int j = 0; j < d - r; ++j)
              if (T[j] == e.first) {
                in = true;
                break;
              }
            if (!in) A.push_back(e.first);
          }
        } else {
          for (auto& e : Rv) {
            if (int(A.size()) >= h) break;
            bool in = false;
            for (int j = 0; j < h; ++j)
              if (T[j] == e.first) {
                in = true;
                break;
              }
            if (!in) A.push_back(e.first);
          }
        }
        const int a = int(A.size());
        for (int i = 0; i < a; ++i) row[h + i] = A[i];
        for (int i = 0; h + a + i < d; ++i) row[h + a + i] = T[h + i];
        out_deg[u] = uint32_t(d);
      }
    }
    if (rev_cnt || rev_keys) {
      for (int64_t u = 0; u < N; ++u) {
        uint32_t nr = roff[u + 1] - roff[u];
        if (rev_cnt) rev_cnt[u] = nr;
        if (rev_keys)
          for (uint32_t k = 0; k < nr; ++k) rev_keys[u * FD + k] = rkeys[roff[u] + k];
      }
    }
  }
  if (fwd_deg) std::memcpy(fwd_deg, fdeg.data(), sizeof(uint32_t) * N);
  if (fwd_ids) std::memcpy(fwd_ids, fids.data(), sizeof(uint32_t) * N * FD);
  if (fwd_d) std::memcpy(fwd_d, fdd.data(), sizeof(float) * N * FD);
  if (counters) {
    counters[0] = pc.collect_evals;
    counters[1] = pc.filter_evals;
    counters[2] = pc.reverse_evals;
    counters[3] = pc.adj_rows;
    counters[4] = pc.ep_evals;
  }
  return 0;
}

// ---------------------------------------------------------------------------------------
// Index writers (common.cu:151-248; spec A.11).  deg/ids: out_deg / out_ids of or_prune.
// Returns 0 on success.
// ---------------------------------------------------------------------------------------
static int write_all(FILE* f, const void* p, size_t n) {
  return fwrite(p, 1, n, f) == n ? 0 : 1;
}

// kind: 0 NSG, 1 Vamana, 2 NSSG, 3 DPG, 4 CAGRA.  nssg_eps: the 10 NSSG entry ids.
int or_store(const char* path, int kind, int64_t N, int FD, const uint32_t* deg,
             const uint32_t* ids, uint32_t ep, const uint32_t* nssg_eps) {
  FILE* f = fopen(path, "wb");
  if (!f) return 1;
  int rc = 0;
  const uint32_t width = uint32_t(FD);
  if (kind == 0 || kind == 2) {
    rc |= write_all(f, &width, 4);
    if (kind == 0) {
      rc |= write_all(f, &ep, 4);
    } else {
      const uint32_t n_ep = 10;
      rc |= write_all(f, &n_ep, 4);
      rc |= write_all(f, nssg_eps, 40);
    }
    for (int64_t u = 0; u < N; ++u) {
      rc |= write_all(f, &deg[u], 4);
      rc |= write_all(f, ids + u * FD, 4ull * deg[u]);
    }
  } else if (kind == 1) {
    uint64_t size = 24;
    uint32_t maxd = 0;
    for (int64_t u = 0; u < N; ++u) {
      size += 4ull * (deg[u] + 1);
      maxd = std::max(maxd, deg[u]);
    }
    const uint64_t frozen = 0;
    rc |= write_all(f, &size, 8);
    rc |= write_all(f, &maxd, 4);
    rc |= write_all(f, &ep, 4);
    rc |= write_all(f, &frozen, 8);
    for (int64_t u = 0; u < N; ++u) {
      rc |= write_all(f, &deg[u], 4);
      rc |= write_all(f, ids + u * FD, 4ull * deg[u]);
    }
  } else if (kind == 3) {
    const char magic[8] = {'K', 'N', 'N', 'G', 'R', 'A', 'P', 'H'};
    const uint32_t vmaj = 2, vmin = 0, n32 = uint32_t(N);
    rc |= write_all(f, magic, 8);
    rc |= write_all(f, &vmaj, 4);
    rc |= write_all(f, &vmin, 4);
    rc |= write_all(f, &n32, 4);
    for (int64_t u = 0; u < N; ++u) {
      rc |= write_all(f, &deg[u], 4);
      rc |= write_all(f, &deg[u], 4);
      rc |= write_all(f, ids + u * FD, 4ull * deg[u]);
    }
  } else if (kind == 4) {
    for (int64_t u = 0; u < N; ++u) {
      rc |= write_all(f, &width, 4);
      rc |= write_all(f, ids + u * FD, 4ull * FD);
    }
  } else {
    rc = 2;
  }
  rc |= fclose(f) != 0;
  return rc;
}

// NSSG entry ids: first 10 of a Fisher-Yates shuffle of 0..N-1 driven by SplitMix64(seed)
// (replaces the unseeded std::random_shuffle of common.cu:201-204).
void or_nssg_eps(int64_t N, uint64_t seed, uint32_t* eps) {
  std::vector<uint32_t> ids(N);
  for (int64_t i = 0; i < N; ++i) ids[i] = uint32_t(i);
  uint64_t s = seed;
  const int64_t m = std::min<int64_t>(10, N);
  for (int64_t i = 0; i < m; ++i) {
    uint64_t j = i + (splitmix64(s) % uint64_t(N - i));
    std::swap(ids[i], ids[j]);
  }
  for (int i = 0; i < 10; ++i) eps[i] = i < m ? ids[i] : 0;
}

// ---------------------------------------------------------------------------------------
// GNN-Descent restatement (Tagore_src.cu:150-1206 driven by pybind.cu:101-125).
// Deterministic spec; identical on the GPU in exact mode:
//   init: per node u a SplitMix64 stream stream_seed(seed,u); draw ids bounded(x,N),
//         rejecting u and repeats, until K ids; canonical distances; sort by (d,id).
//   phase 1 (iters/2 times):
//     sample: new[u] = first S unvisited entries, old[u] = first S visited entries
//             (closed form of the three scan loops, Tagore_src.cu:536-567); the new ones
//             are marked visited; push (d,u) to rnew[v] / rold[v].
//     reverse sample: rnew[v], rold[v] = S smallest pushes; new'[v] = new[v] ++ (rnew[v]
//             minus new[v]), old'[v] likewise (Tagore_src.cu:603-674).
//     local join (Tagore_src.cu:684-1033): A = new'[u], B = old'[u]; for every row i of A
//             and 16-wide tile of B the (d,id)-min pair (excluding equal ids) is pushed to
//             A[i]; for every column j of B and 16-tall tile of A the min is pushed to B[j];
//             for every row i of A and 16-wide tile of A (j != i) the min is pushed to A[i].
//             A push survives only if d < worst(dest) (the list's K-th distance).
//     merge (Tagore_src.cu:1043-1161): drop pushes that repeat, name u, are already
//             listed or are not below worst; keep the K smallest of list u the rest
//             (the reference's 128-slot cap is a capacity limit; since the merge keeps
//             only K <= 128 entries, an uncapped candidate set gives the exact result);
//             inserted entries are unvisited.
//   reverse build (Tagore_src.cu:1187-1206): R[v] = K smallest (d,u) over u listing v;
//             all flags cleared.
//   phase 2 (iters/2 times, reading a snapshot G0,R0; Tagore_src.cu:178-403):
//     leaders = first min(4,#unvisited) unvisited entries of u (marked visited);
//     candidates = for each leader v: G0[v][0..min(64,K)) then R0[v][0..min(64,|R0[v]|));
//     drop repeats, u, ids in G0[u]; keep d < worst; merge; push (d,u) to each newly
//     inserted w; R[w] = K smallest of R0[w] u pushes.
//   final reverse merge (merge_reverse_plus, Tagore_src.cu:420-512): candidates R[u]
//     with d < worst, not listed, not u; merge.
// Outputs G (N*K ids), GD (N*K dists).  Returns 0 on success.
// ---------------------------------------------------------------------------------------
// or_gnn_descent_from: the same iterations started from given lists G0/GD0 (rows sorted by
// (d,id), no self, no repeats, canonical distances; all entries unvisited).  iters == 0
// returns the initial lists unchanged (no reverse merge).
int or_gnn_descent_from(const float* X, int64_t N, int D, int K, int iters, uint64_t seed,
                        const uint32_t* G0, const float* GD0, uint32_t* G, float* GD,
                        uint64_t* counters);

int or_gnn_descent(const float* X, int64_t N, int D, int K, int iters, uint64_t seed,
                   uint32_t* G, float* GD, uint64_t* counters) {
  return or_gnn_descent_from(X, N, D, K, iters, seed, nullptr, nullptr, G, GD, counters);
}

int or_gnn_descent_from(const float* X, int64_t N, int D, int K, int iters, uint64_t seed,
                        const uint32_t* Ginit, const float* GDinit, uint32_t* G, float* GD,
                        uint64_t* counters) {
  const int S = 32, TILE = 16, P2_LEADERS = 4, P2_TAKE = 64;
  if (K < 1 || N <= K) return 1;
  const int nthreads = omp_get_max_threads();
  std::vector<uint8_t> V(size_t(N) * K, 0);
  uint64_t evals = 0;
  if (Ginit) {
    std::copy(Ginit, Ginit + size_t(N) * K, G);
    std::copy(GDinit, GDinit + size_t(N) * K, GD);
  } else {
  // ---- init ----
#pragma omp parallel
  {
    std::vector<Ent> row(K);
#pragma omp for schedule(static)
    for (int64_t u = 0; u < N; ++u) {
      uint64_t st = stream_seed(seed, uint64_t(u));
      int cnt = 0;
      while (cnt < K) {
        uint32_t c = bounded(splitmix64(st), uint32_t(N));
        if (c == uint32_t(u)) continue;
        bool dup = false;
        for (int j = 0; j < cnt; ++j)
          if (row[j].id == c) {
            dup = true;
            break;
          }
        if (dup) continue;
        row[cnt++].id = c;
      }
      for (int j = 0; j < K; ++j) row[j].d = cdist(X + u * D, X + uint64_t(row[j].id) * D, D);
      std::sort(row.begin(), row.end(), ent_lt);
      for (int j = 0; j < K; ++j) {
        G[u * K + j] = row[j].id;
        GD[u * K + j] = row[j].d;
      }
    }
  }
  evals += uint64_t(N) * K;
  }
  if (iters == 0) {
    if (counters) counters[0] = evals;
    return 0;
  }

  auto merge_into = [&](int64_t u, std::vector<Ent>& cand, uint32_t* g, float* gd,
                        uint8_t* vis) {
    // cand: sorted, unique, all < worst, none already listed.  Keep K smallest.
    std::vector<Ent> out;
    std::vector<uint8_t> ov;
    out.reserve(K);
    int a = 0, b = 0;
    while (int(out.size()) < K) {
      bool takeA;
      if (a >= K) takeA = false;
      else if (b >= int(cand.size())) takeA = true;
      else takeA = ent_lt(Ent{gd[a], g[a]}, cand[b]);
      if (takeA) {
        out.push_back({gd[a], g[a]});
        ov.push_back(vis[a]);
        ++a;
      } else {
        out.push_back(cand[b]);
        ov.push_back(0);
        ++b;
      }
    }
    (void)u;
    for (int j = 0; j < K; ++j) {
      g[j] = out[j].id;
      gd[j] = out[j].d;
      vis[j] = ov[j];
    }
  };
  // Candidate filter shared by every merge: unique keys, not self, not listed, d < worst.
  auto clean_cands = [&](int64_t u, std::vector<Ent>& c, const uint32_t* g, const float* gd) {
    std::sort(c.begin(), c.end(), ent_lt);
    c.erase(std::unique(c.begin(), c.end(), ent_eq), c.end());
    const float worst = gd[K - 1];
    std::vector<Ent> keep;
    for (auto& e : c) {
      if (e.id == uint32_t(u) || !(e.d < worst)) continue;
      bool listed = false;
      for (int j = 0; j < K; ++j)
        if (g[j] == e.id) {
          listed = true;
          break;
        }
      if (!listed) keep.push_back(e);
    }
    c.swap(keep);
  };

  const int it1 = iters / 2, it2 = iters / 2;
  std::vector<uint32_t> hyb(size_t(N) * 4 * S);  // new' [0,2S), old' [2S,4S)
  std::vector<uint32_t> na(N), nb(N);
  for (int it = 0; it < it1; ++it) {
    // ---- sample ----
    PushList pn(nthreads), po(nthreads);
    std::vector<uint32_t> nnew(N), nold(N);
#pragma omp parallel
    {
      const int tid = omp_get_thread_num();
#pragma omp for schedule(static)
      for (int64_t u = 0; u < N; ++u) {
        uint32_t* h = hyb.data() + u * 4 * S;
        int cn = 0, co = 0;
        for (int j = 0; j < K; ++j) {
          uint8_t& f = V[u * K + j];
          if (!f) {
            if (cn < S) {
              h[cn++] = G[u * K + j];
              f = 1;
              pn.push(tid, G[u * K + j], ent_key(GD[u * K + j], uint32_t(u)));
            }
          } else if (co < S) {
            h[2 * S + co++] = G[u * K + j];
            po.push(tid, G[u * K + j], ent_key(GD[u * K + j], uint32_t(u)));
          }
        }
        nnew[u] = cn;
        nold[u] = co;
      }
    }
    std::vector<uint32_t> offn, offo;
    std::vector<uint64_t> kn, ko;
    pn.reduce(uint32_t(N), S, offn, kn);
    po.reduce(uint32_t(N), S, offo, ko);
    // ---- reverse sample ----
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < N; ++v) {
      uint32_t* h = hyb.data() + v * 4 * S;
      uint32_t cn = nnew[v], co = nold[v];
      const uint32_t n0 = cn, o0 = co;
      for (uint32_t k = offn[v]; k < offn[v + 1]; ++k) {
        uint32_t w = key_id(kn[k]);
        bool in = false;
        for (uint32_t j = 0; j < n0; ++j)
          if (h[j] == w) {
            in = true;
            break;
          }
        if (!in) h[cn++] = w;
      }
      for (uint32_t k = offo[v]; k < offo[v + 1]; ++k) {
        uint32_t w = key_id(ko[k]);
        bool in = false;
        for (uint32_t j = 0; j < o0; ++j)
          if (h[2 * S + j] == w) {
            in = true;
            break;
          }
        if (!in) h[2 * S + co++] = w;
      }
      na[v] = cn;
      nb[v] = co;
    }
    // ---- local join ----
    PushList pc(nthreads);
    std::vector<uint64_t> ev(nthreads, 0);
#pragma omp parallel
    {
      const int tid = omp_get_thread_num();
      std::vector<float> MAB(size_t(2 * S) * 2 * S), MAA(size_t(2 * S) * 2 * S);
#pragma omp for schedule(dynamic, 16)
      for (int64_t u = 0; u < N; ++u) {
        const uint32_t* A = hyb.data() + u * 4 * S;
        const uint32_t* Bv = A + 2 * S;
        const int nA = int(na[u]), nB = int(nb[u]);
        for (int i = 0; i < nA; ++i)
          for (int j = 0; j < nB; ++j)
            MAB[i * 2 * S + j] = (A[i] == Bv[j])
                                     ? FINF
                                     : cdist(X + uint64_t(A[i]) * D, X + uint64_t(Bv[j]) * D, D);
        for (int i = 0; i < nA; ++i) {
          MAA[i * 2 * S + i] = FINF;
          for (int j = i + 1; j < nA; ++j) {
            float dd = cdist(X + uint64_t(A[i]) * D, X + uint64_t(A[j]) * D, D);
            MAA[i * 2 * S + j] = dd;
            MAA[j * 2 * S + i] = dd;
          }
        }
        ev[tid] += uint64_t(nA) * nB + uint64_t(nA) * (nA - 1) / 2;
        auto push_if = [&](uint32_t dest, float dd, uint32_t id) {
          if (dd < GD[uint64_t(dest) * K + K - 1]) pc.push(tid, dest, ent_key(dd, id));
        };
        for (int i = 0; i < nA; ++i)
          for (int t0 = 0; t0 < nB; t0 += TILE) {
            Ent best{FINF, INV};
            for (int j = t0; j < std::min(nB, t0 + TILE); ++j) {
              Ent e{MAB[i * 2 * S + j], Bv[j]};
              if (e.d < FINF && ent_lt(e, best)) best = e;
            }
            if (best.id != INV) push_if(A[i], best.d, best.id);
          }
        for (int j = 0; j < nB; ++j)
          for (int s0 = 0; s0 < nA; s0 += TILE) {
            Ent best{FINF, INV};
            for (int i = s0; i < std::min(nA, s0 + TILE); ++i) {
              Ent e{MAB[i * 2 * S + j], A[i]};
              if (e.d < FINF && ent_lt(e, best)) best = e;
            }
            if (best.id != INV) push_if(Bv[j], best.d, best.id);
          }
        for (int i = 0; i < nA; ++i)
          for (int t0 = 0; t0 < nA; t0 += TILE) {
            Ent best{FINF, INV};
            for (int j = t0; j < std::min(nA, t0 + TILE); ++j) {
              Ent e{MAA[i * 2 * S + j], A[j]};
              if (e.d < FINF && ent_lt(e, best)) best = e;
            }
            if (best.id != INV) push_if(A[i], best.d, best.id);
          }
      }
    }
    for (auto e : ev) evals += e;
    std::vector<uint32_t> offc;
    std::vector<uint64_t> kc;
    pc.reduce(uint32_t(N), 0, offc, kc);
    // ---- merge ----
#pragma omp parallel
    {
      std::vector<Ent> cand;
#pragma omp for schedule(dynamic, 64)
      for (int64_t v = 0; v < N; ++v) {
        cand.clear();
        for (uint32_t k = offc[v]; k < offc[v + 1]; ++k)
          cand.push_back({key_d(kc[k]), key_id(kc[k])});
        clean_cands(v, cand, G + v * K, GD + v * K);
        if (!cand.empty()) merge_into(v, cand, G + v * K, GD + v * K, V.data() + v * K);
      }
    }
  }
  // ---- reverse build + flag reset ----
  std::vector<uint32_t> roff;
  std::vector<uint64_t> rk;
  {
    PushList pr(nthreads);
#pragma omp parallel
    {
      const int tid = omp_get_thread_num();
#pragma omp for schedule(static)
      for (int64_t u = 0; u < N; ++u)
        for (int j = 0; j < K; ++j)
          pr.push(tid, G[u * K + j], ent_key(GD[u * K + j], uint32_t(u)));
    }
    pr.reduce(uint32_t(N), uint32_t(K), roff, rk);
  }
  std::fill(V.begin(), V.end(), 0);
  // ---- phase 2 ----
  std::vector<uint32_t> G0(size_t(N) * K);
  std::vector<float> GD0(size_t(N) * K);
  for (int it = 0; it < it2; ++it) {
    std::memcpy(G0.data(), G, sizeof(uint32_t) * N * K);
    std::memcpy(GD0.data(), GD, sizeof(float) * N * K);
    PushList pr(nthreads);
    std::vector<uint64_t> ev(nthreads, 0);
#pragma omp parallel
    {
      const int tid = omp_get_thread_num();
      std::vector<uint32_t> ids;
      std::vector<Ent> cand;
#pragma omp for schedule(dynamic, 16)
      for (int64_t u = 0; u < N; ++u) {
        const uint32_t* g0 = G0.data() + u * K;
        const float* gd0 = GD0.data() + u * K;
        uint8_t* vis = V.data() + u * K;
        ids.clear();
        int nl = 0;
        for (int j = 0; j < K && nl < P2_LEADERS; ++j)
          if (!vis[j]) {
            vis[j] = 1;
            ++nl;
            const uint32_t v = g0[j];
            const int nf = std::min(P2_TAKE, K);
            for (int k = 0; k < nf; ++k) ids.push_back(G0[uint64_t(v) * K + k]);
            const uint32_t nr = std::min<uint32_t>(P2_TAKE, roff[v + 1] - roff[v]);
            for (uint32_t k = 0; k < nr; ++k) ids.push_back(key_id(rk[roff[v] + k]));
          }
        std::sort(ids.begin(), ids.end());
        ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
        cand.clear();
        const float worst = gd0[K - 1];
        for (uint32_t c : ids) {
          if (c == uint32_t(u)) continue;
          bool listed = false;
          for (int j = 0; j < K; ++j)
            if (g0[j] == c) {
              listed = true;
              break;
            }
          if (listed) continue;
          float dd = cdist(X + u * D, X + uint64_t(c) * D, D);
          ev[tid]++;
          if (dd < worst) cand.push_back({dd, c});
        }
        std::sort(cand.begin(), cand.end(), ent_lt);
        // G/GD rows equal G0/GD0 rows here (each row is written only by its own node).
        if (!cand.empty()) {
          merge_into(u, cand, G + u * K, GD + u * K, vis);
          for (int j = 0; j < K; ++j) {
            bool was = false;
            for (int k = 0; k < K; ++k)
              if (g0[k] == G[u * K + j]) {
                was = true;
                break;
              }
            if (!was) pr.push(tid, G[u * K + j], ent_key(GD[u * K + j], uint32_t(u)));
          }
        }
      }
    }
    for (auto e : ev) evals += e;
    // R[w] = K smallest of R0[w] u pushes (identical keys merged).
    for (int64_t w = 0; w < N; ++w)
      for (uint32_t k = roff[w]; k < roff[w + 1]; ++k) pr.push(0, uint32_t(w), rk[k]);
    std::vector<uint32_t> noff;
    std::vector<uint64_t> nk;
    pr.reduce(uint32_t(N), 0, noff, nk);
    // unique + cap K
    std::vector<uint32_t> off2(N + 1, 0);
    std::vector<uint64_t> k2;
    k2.reserve(nk.size());
    for (int64_t w = 0; w < N; ++w) {
      uint32_t n = 0;
      uint64_t prev = ~0ull;
      for (uint32_t k = noff[w]; k < noff[w + 1] && n < uint32_t(K); ++k) {
        if (nk[k] == prev) continue;
        prev = nk[k];
        k2.push_back(nk[k]);
        ++n;
      }
      off2[w + 1] = off2[w] + n;
    }
    roff.swap(off2);
    rk.swap(k2);
  }
  // ---- final reverse merge ----
#pragma omp parallel
  {
    std::vector<Ent> cand;
#pragma omp for schedule(dynamic, 64)
    for (int64_t u = 0; u < N; ++u) {
      cand.clear();
      for (uint32_t k = roff[u]; k < roff[u + 1]; ++k)
        cand.push_back({key_d(rk[k]), key_id(rk[k])});
      clean_cands(u, cand, G + u * K, GD + u * K);
      if (!cand.empty()) merge_into(u, cand, G + u * K, GD + u * K, V.data() + u * K);
    }
  }
  if (counters) counters[0] = evals;
  return 0;
}

// Direct access to the Filter step for rule-level tests: pool (ids, d) of length L is
// sorted and deduplicated (as every caller does), then filtered.  Writes the kept ids/d
// and returns their count.
int or_filter_pool(const float* X, int D, const uint32_t* pid, const float* pd, int L,
                   uint32_t u, int FD, int rule, float th, uint32_t* out_ids, float* out_d) {
  std::vector<Ent> P(L), out;
  for (int i = 0; i < L; ++i) P[i] = {pd[i], pid[i]};
  sort_dedup_pool(P);
  filter_pool(P, u, FD, rule, th, X, D, out, nullptr);
  for (size_t k = 0; k < out.size(); ++k) {
    out_ids[k] = out[k].id;
    out_d[k] = out[k].d;
  }
  return int(out.size());
}

// Direct access to the path Collect for tests: returns the sorted, deduplicated pool of
// node u (length K + 480) and the ring counter.
uint64_t or_collect_path(const float* X, int D, const uint32_t* G, const float* GD, int K,
                         int TOPM, uint32_t ep, uint32_t u, uint32_t* pool_ids,
                         float* pool_d) {
  const float* xu = X + uint64_t(u) * D;
  auto dfn = [&](uint32_t y) { return cdist(xu, X + uint64_t(y) * D, D); };
  std::vector<Ent> ring(RING, Ent{FINF, INV}), B, P;
  std::vector<uint8_t> ex;
  uint64_t c = 0;
  auto on_eval = [&](const Ent& e) { ring[c % RING] = e; ++c; };
  beam_search(G, K, TOPM, ep, PATH_EXPANSIONS, dfn, on_eval, B, ex, nullptr, nullptr);
  for (int j = 0; j < K; ++j) P.push_back({GD[uint64_t(u) * K + j], G[uint64_t(u) * K + j]});
  for (auto& e : ring) P.push_back(e);
  sort_dedup_pool(P);
  for (size_t i = 0; i < P.size(); ++i) {
    pool_ids[i] = P[i].id;
    pool_d[i] = P[i].d;
  }
  return c;
}

// Direct access to the two-hop Collect for tests: the sorted, deduplicated pool of node u
// (length K + 480): its own list (ids and distances from GD) and the 480 two-hop slots.
void or_collect_2hop(const float* X, int D, const uint32_t* G, const float* GD, int K,
                     uint32_t u, uint32_t* pool_ids, float* pool_d) {
  const float* xu = X + uint64_t(u) * D;
  auto dfn = [&](uint32_t y) { return cdist(xu, X + uint64_t(y) * D, D); };
  std::vector<Ent> P;
  for (int j = 0; j < K; ++j) P.push_back({GD[uint64_t(u) * K + j], G[uint64_t(u) * K + j]});
  collect_2hop_pool(G, K, int64_t(u), dfn, P);
  sort_dedup_pool(P);
  for (size_t i = 0; i < P.size(); ++i) {
    pool_ids[i] = P[i].id;
    pool_d[i] = P[i].d;
  }
}

// Direct access to the reverse phase for tests: forward rows (fdeg [N], fids/fdd [N x FD])
// and a list of Store pushes (dst, src, d) -- src is pushed to dst with distance d -- give
// the final rows exactly as or_prune's reverse phase does.  Returns the rule evaluations.
uint64_t or_reverse(const float* X, int64_t N, int D, int K, int FD, int rev_rule, float th,
                    const uint32_t* fdeg, const uint32_t* fids, const float* fdd,
                    int64_t npush, const uint32_t* dst, const uint32_t* src, const float* d,
                    uint32_t* out_deg, uint32_t* out_ids) {
  PushList pushes(1);
  for (int64_t i = 0; i < npush; ++i) pushes.push(0, dst[i], ent_key(d[i], src[i]));
  return reverse_phase(X, N, D, K, FD, rev_rule, th, fdeg, fids, fdd, pushes, out_deg,
                       out_ids, nullptr, nullptr);
}

// ---------------------------------------------------------------------------------------
// Beam search over a pruned index (get_neighbors, common.cu:65-127), restated for the
// variable-degree rows (deg[u], ids[u*FD + j]) of a pruned graph:
//   * initial ids: the first min(L, deg(ep)) neighbours of ep, then draws
//     bounded(stream_seed(qs, t), N), t = 0, 1, ... with qs = seed ^ ((q+1) * 0xA0761D6478BD642F)
//     (the reference: rand() % N), skipping visited ids, until min(L, N) ids or 4096 draws;
//     every initial id is marked visited;
//   * the pool: those ids with canonical distances, sorted by (d, id), all unchecked;
//   * k = 0; repeat: k = first unchecked position >= k (stop if none); mark it checked and
//     expand its node n: every neighbour not yet visited is marked visited and evaluated;
//     candidates below the pool's last entry enter it one at a time, the last entry leaving
//     when the pool holds L (InsertIntoPool, common.cu:25-52, with (d, id) order); nk = the
//     smallest insertion position (L if none); k = nk <= k ? nk : k + 1.
//   * output: the first k entries (ids, canonical distances); INV / +inf past the pool.
// The reference's compare() returns the square root of an FP32 running sum; here the
// canonical squared distance orders the pool.  counters: [0] evaluations, [1] expansions.
// ---------------------------------------------------------------------------------------
int or_search(const float* X, int64_t N, int D, const uint32_t* deg, const uint32_t* ids,
              int FD, uint32_t ep, const float* Q, int64_t nq, int L, int k, uint64_t seed,
              uint32_t* out_ids, float* out_d, uint64_t* counters) {
  uint64_t evals = 0, expansions = 0;
  std::vector<uint8_t> vis(N);
  for (int64_t q = 0; q < nq; ++q) {
    std::fill(vis.begin(), vis.end(), uint8_t(0));
    const float* xq = Q + q * D;
    const int64_t target = std::min<int64_t>(L, N);
    std::vector<uint32_t> init;
    const int d0 = std::min<int>(int(deg[ep]), L);
    for (int j = 0; j < d0; ++j) {
      const uint32_t y = ids[uint64_t(ep) * FD + j];
      init.push_back(y);
      vis[y] = 1;
    }
    const uint64_t qs = seed ^ ((uint64_t(q) + 1) * 0xA0761D6478BD642FULL);
    for (uint64_t t = 0; int64_t(init.size()) < target && t < 4096; ++t) {
      const uint32_t y = bounded(stream_seed(qs, t), uint32_t(N));
      if (vis[y]) continue;
      vis[y] = 1;
      init.push_back(y);
    }
    std::vector<Ent> pool;
    std::vector<uint8_t> checked;
    for (uint32_t y : init) pool.push_back({cdist(xq, X + uint64_t(y) * D, D), y});
    evals += init.size();
    std::sort(pool.begin(), pool.end(), ent_lt);
    checked.assign(pool.size(), 0);
    int kk = 0;
    for (;;) {
      while (kk < int(pool.size()) && checked[kk]) ++kk;
      if (kk >= int(pool.size())) break;
      checked[kk] = 1;
      const uint32_t n = pool[kk].id;
      ++expansions;
      int nk = L;
      for (uint32_t j = 0; j < deg[n]; ++j) {
        const uint32_t y = ids[uint64_t(n) * FD + j];
        if (vis[y]) continue;
        vis[y] = 1;
        const Ent e{cdist(xq, X + uint64_t(y) * D, D), y};
        ++evals;
        if (int(pool.size()) == L && !ent_lt(e, pool.back())) continue;
        const int r = int(std::lower_bound(pool.begin(), pool.end(), e, ent_lt) - pool.begin());
        pool.insert(pool.begin() + r, e);
        checked.insert(checked.begin() + r, 0);
        if (int(pool.size()) > L) {
          pool.pop_back();
          checked.pop_back();
        }
        nk = std::min(nk, r);
      }
      kk = nk <= kk ? nk : kk + 1;
    }
    for (int j = 0; j < k; ++j) {
      out_ids[q * k + j] = j < int(pool.size()) ? pool[j].id : INV;
      out_d[q * k + j] = j < int(pool.size()) ? pool[j].d : FINF;
    }
  }
  if (counters) {
    counters[0] = evals;
    counters[1] = expansions;
  }
  return 0;
}

// ---------------------------------------------------------------------------------------
// Learned-init encoder (no counterpart in the reference, whose lists start random:
// initialize_graph, Tagore_src.cu:158-166).  FP64 restatement of the HIP forward/backward
// with the same BF16 rounding points when round != 0:
//   xn = bf16((x - mu) * s); h = bf16(relu(bf16(W1) xn + b1)); z = bf16(W2) h + b2;
//   L = 1/(B P) sum_{b,j} |z_b - z_p(b,j)|^2 + 1/(B (B-1)) sum_{b != c} relu(m - |z_b - z_c|^2)
//     (rows [0,B) anchors, row B + b P + j positive j of anchor b; a pair is active iff
//      |z_b - z_c|^2 < m);
//   dz -> bf16; [dW2 | db2] = dz^T [h | 1]; dh = bf16((dz bf16(W2)) * [h > 0]);
//   [dW1 | db1] = dh^T [xn | 1].
// With round == 0 every quantity is plain FP64, the form the finite-difference test pins.
// Parameter layout: [256][D+1] (W1 row | b1) then [32][257] (W2 row | b2).
// ---------------------------------------------------------------------------------------
static double bf16_round(double x, int round) {
  if (!round) return x;
  float f = float(x);
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u = (u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u;
  std::memcpy(&f, &u, 4);
  return double(f);
}

int or_mlp_eval(const float* X, int64_t M, int D, int B, int P, const float* mu, float scale,
                const double* prm, double margin, int round, double* loss, double* grad,
                double* codes) {
  const int H = 256, C = 32;
  if (M != int64_t(B) * (1 + P) || B < 2) return 1;
  const size_t n1 = size_t(H) * (D + 1);
  std::vector<double> W1(size_t(H) * D), W2(size_t(C) * H);
  for (int j = 0; j < H; ++j)
    for (int i = 0; i < D; ++i) W1[size_t(j) * D + i] = bf16_round(prm[size_t(j) * (D + 1) + i], round);
  for (int k = 0; k < C; ++k)
    for (int j = 0; j < H; ++j) W2[size_t(k) * H + j] = bf16_round(prm[n1 + size_t(k) * (H + 1) + j], round);
  std::vector<double> xn(size_t(M) * D), h(size_t(M) * H), z(size_t(M) * C);
  for (int64_t m = 0; m < M; ++m)
    for (int i = 0; i < D; ++i)
      xn[m * D + i] = bf16_round(round ? double(float((X[m * D + i] - mu[i]) * scale))
                                       : (double(X[m * D + i]) - mu[i]) * double(scale),
                                 round);
#pragma omp parallel for schedule(static)
  for (int64_t m = 0; m < M; ++m) {
    for (int j = 0; j < H; ++j) {
      double a = prm[size_t(j) * (D + 1) + D];
      for (int i = 0; i < D; ++i) a += W1[size_t(j) * D + i] * xn[m * D + i];
      h[m * H + j] = bf16_round(a > 0 ? a : 0.0, round);
    }
    for (int k = 0; k < C; ++k) {
      double a = prm[n1 + size_t(k) * (H + 1) + H];
      for (int j = 0; j < H; ++j) a += W2[size_t(k) * H + j] * h[m * H + j];
      z[m * C + k] = a;
    }
  }
  std::vector<double> dz(size_t(M) * C, 0.0);
  const double cpos = 2.0 / (double(B) * P), cneg = 4.0 / (double(B) * (B - 1));
  double lpos = 0, lneg = 0;
  for (int b = 0; b < B; ++b) {
    for (int j = 0; j < P; ++j) {
      const int64_t r = B + int64_t(b) * P + j;
      for (int k = 0; k < C; ++k) {
        const double e = z[b * C + k] - z[r * C + k];
        lpos += e * e;
        dz[b * C + k] += cpos * e;
        dz[r * C + k] += -cpos * e;
      }
    }
    for (int c = 0; c < B; ++c) {
      if (c == b) continue;
      double d2 = 0;
      for (int k = 0; k < C; ++k) {
        const double e = z[b * C + k] - z[c * C + k];
        d2 += e * e;
      }
      if (d2 < margin) {
        lneg += margin - d2;
        for (int k = 0; k < C; ++k) dz[b * C + k] -= cneg * (z[b * C + k] - z[c * C + k]);
      }
    }
  }
  *loss = lpos / (double(B) * P) + lneg / (double(B) * (B - 1));
  for (auto& v : dz) v = bf16_round(v, round);
  // [dW2 | db2]
  for (int k = 0; k < C; ++k) {
    for (int j = 0; j <= H; ++j) {
      double a = 0;
      for (int64_t m = 0; m < M; ++m) a += dz[m * C + k] * (j < H ? h[m * H + j] : 1.0);
      grad[n1 + size_t(k) * (H + 1) + j] = a;
    }
  }
  // dh, then [dW1 | db1]
  std::vector<double> dh(size_t(M) * H);
  for (int64_t m = 0; m < M; ++m)
    for (int j = 0; j < H; ++j) {
      double a = 0;
      for (int k = 0; k < C; ++k) a += dz[m * C + k] * W2[size_t(k) * H + j];
      dh[m * H + j] = bf16_round(h[m * H + j] > 0 ? a : 0.0, round);
    }
#pragma omp parallel for schedule(static)
  for (int j = 0; j < H; ++j)
    for (int i = 0; i <= D; ++i) {
      double a = 0;
      for (int64_t m = 0; m < M; ++m) a += dh[m * H + j] * (i < D ? xn[m * D + i] : 1.0);
      grad[size_t(j) * (D + 1) + i] = a;
    }
  if (codes) std::copy(z.begin(), z.end(), codes);
  return 0;
}

// Adam step t (Kingma & Ba 2015, Algorithm 1) in FP64: beta1 0.9, beta2 0.999, eps 1e-8,
// each constant taken at its FP32 value (the values the HIP kernel holds).
void or_mlp_adam(int64_t n, int t, double lr, const double* g, double* p, double* m, double* v) {
  const double b1 = double(0.9f), b2 = double(0.999f), eps = double(1e-8f);
  const double bc1 = 1.0 - std::pow(b1, t), bc2 = 1.0 - std::pow(b2, t);
  for (int64_t i = 0; i < n; ++i) {
    m[i] = b1 * m[i] + (1 - b1) * g[i];
    v[i] = b2 * v[i] + (1 - b2) * g[i] * g[i];
    p[i] -= lr * (m[i] / bc1) / (std::sqrt(v[i] / bc2) + eps);
  }
}

int or_num_threads(void) { return omp_get_max_threads(); }
void or_set_threads(int n) { omp_set_num_threads(n); }

}  // extern "C"
