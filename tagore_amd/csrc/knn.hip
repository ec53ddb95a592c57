// knn.hip -- GNN-Descent kNN-graph construction kernels for gfx950 (exact FP32 mode).
//
// Deterministic restatement of Tagore's two-phase GNN-Descent (Tagore_src.cu:150-1206 as
// driven by pybind.cu:101-125); the CPU oracle (or_gnn_descent) states the same spec, so
// in exact mode the graph produced here is bit-identical to it:
//   init                 initialize_graph           Tagore_src.cu:158-166
//   sample / rsample     nn_descent_opt_sample      Tagore_src.cu:531-594, 603-674
//   local join           nn_descent_opt_cal         Tagore_src.cu:684-1033
//   merge                nn_descent_opt_merge       Tagore_src.cu:1043-1161
//   reverse build        do_reverse_graph           Tagore_src.cu:1187-1206
//   phase 2              sample_kernel6             Tagore_src.cu:178-403
//   final reverse merge  merge_reverse_plus         Tagore_src.cu:420-512
// Candidate and reverse lists go through PushBufs (order-independent reductions) instead
// of the reference's atomic slot counters.
#include "common.hpp"
#include "kernels.hpp"

namespace tg {

constexpr int S = 32;        // SAMPLE
constexpr int TILE = 16;     // local-join tile edge
constexpr int HROW = 128;    // sample-list row: new' [0,64), old' [64,128)
constexpr int LEADERS = 4;   // phase-2 expansions per node
constexpr int TAKE = 64;     // forward / reverse entries taken per leader

// ------------------------------------------------------------------------------- init
// The i-th draw of node u's SplitMix64 stream is a pure function of (stream start, i), so a
// wave evaluates 64 draws at once and accepts them in stream order: a draw is taken unless it
// names u, repeats an accepted id, or repeats an earlier draw of the same batch (an earlier
// equal draw is either accepted or rejected for a reason that applies to this one too).
__global__ __launch_bounds__(64) void k_init_ids(KnnArgs a) {
  __shared__ uint32_t acc[128];
  __shared__ uint32_t first[TG_WAVE];
  const uint32_t u = blockIdx.x;
  if (u >= a.N) return;
  const int lane = threadIdx.x, K = a.K;
  // a segment draws inside [lo, lo + n) from the stream of its own seed and local row index,
  // exactly as a handle holding only that segment would (ids shifted by lo)
  uint32_t lo = 0, n = a.N;
  uint64_t seed = a.seed;
  if (a.seg_of) {
    const uint32_t c = a.seg_of[u];
    lo = a.seg_lo[c];
    n = a.seg_lo[c + 1] - lo;
    seed = segment_seed(a.seed, a.seg_key[c]);
  }
  const uint64_t st0 = sm_stream(seed, u - lo);
  int cnt = 0;
  for (uint64_t base = 0; cnt < K; base += TG_WAVE) {
    uint64_t s = st0 + (base + lane) * 0x9E3779B97F4A7C15ull;
    const uint32_t c = lo + sm_bounded(sm_next(s), n);
    bool ok = c != u;
    for (int j = 0; j < cnt && ok; ++j) ok = acc[j] != c;
    // a draw repeating one in a lower lane is rejected: sort (id, lane), keep each id's first
    uint64_t kv[1] = {(uint64_t(c) << 32) | uint32_t(lane)};
    reg_bitonic<1>(kv);
    const uint32_t prev = uint32_t(__shfl(int(kv[0] >> 32), (lane + 63) & 63, TG_WAVE));
    first[uint32_t(kv[0])] = (lane == 0 || prev != uint32_t(kv[0] >> 32)) ? 1u : 0u;
    __syncthreads();
    ok = ok && first[lane] != 0;
    const uint64_t m = ballot(ok);
    const int rk = cnt + mask_rank(m);
    __syncthreads();
    if (ok && rk < K) acc[rk] = c;
    cnt = min(K, cnt + int(__popcll(m)));
    __syncthreads();
  }
  for (int j = lane; j < K; j += TG_WAVE) a.G[size_t(u) * K + j] = acc[j];
}

// Distances of the initial lists and their (d, id) sort.  T: FP32 rows, or the BF16 copy
// (Xr16) when every value is exact in BF16 (the same values, half the bytes).
template <int MC, class T>
__global__ __launch_bounds__(64) void k_init_sort(KnnArgs a) {
  __shared__ uint64_t kk[128];
  __shared__ uint32_t row[128];
  const uint32_t u = row_of_block(a.perm, a.N);
  if (u >= a.N) return;
  const int lane = threadIdx.x, t = lane & 15;
  const int K = a.K, D = a.D, n = pow2_at_least(K);
  const T* X = std::is_same<T, uint8_t>::value
                   ? reinterpret_cast<const T*>(a.X8)
                   : (a.Xr16 ? reinterpret_cast<const T*>(a.Xr16) : reinterpret_cast<const T*>(a.X));
  Row16<MC> xu;
  load_row16<MC>(xu, X + size_t(u) * D, D, t);
  // the list's ids in LDS first, so the distance batches wait for one load, not two
  for (int j = lane; j < n; j += TG_WAVE) {
    kk[j] = TG_KEY_INF;
    if (j < K) row[j] = a.G[size_t(u) * K + j];
  }
  __syncthreads();
  eval_batch<MC>(X, D, xu, K, [&](int q) { return row[q]; },
                 [&](int q, float d) { kk[q] = mk_key(d, row[q]); });
  __syncthreads();
  wave_sort<128>(kk, n);
  for (int j = lane; j < K; j += TG_WAVE) {
    a.G[size_t(u) * K + j] = key_id(kk[j]);
    a.GD[size_t(u) * K + j] = key_d(kk[j]);
    a.V[size_t(u) * K + j] = 0;
  }
  if (a.counters && lane == 0) atomicAdd(&cnt_slot(a.counters)[0], (unsigned long long)K);
}

// ------------------------------------------------------------------------- phase 1
// new[u] = first S unvisited entries (marked visited), old[u] = first S visited entries:
// the closed form of the reference's three scan loops (Tagore_src.cu:536-567).
__global__ __launch_bounds__(64) void k_sample(KnnArgs a) {
  const uint32_t u = blockIdx.x;
  if (u >= a.N) return;
  const int lane = threadIdx.x;
  const int K = a.K;
  int cu = 0, cv = 0;
  for (int base = 0; base < K; base += TG_WAVE) {
    const int j = base + lane;
    const bool valid = j < K;
    const bool vis = valid && a.V[size_t(u) * K + j] != 0;
    const bool unv = valid && !vis;
    const uint64_t mU = ballot(unv), mV = ballot(vis);
    const int rU = cu + mask_rank(mU), rV = cv + mask_rank(mV);
    if (unv && rU < S) {
      const uint32_t id = a.G[size_t(u) * K + j];
      a.H[size_t(u) * HROW + rU] = id;
      a.V[size_t(u) * K + j] = 1;
      pb_push(a.pn, id, mk_key(a.GD[size_t(u) * K + j], u));
    }
    if (vis && rV < S) {
      const uint32_t id = a.G[size_t(u) * K + j];
      a.H[size_t(u) * HROW + 64 + rV] = id;
      pb_push(a.po, id, mk_key(a.GD[size_t(u) * K + j], u));
    }
    cu += __popcll(mU);
    cv += __popcll(mV);
  }
  if (lane == 0) {
    a.nnew[u] = min(cu, S);
    a.nold[u] = min(cv, S);
  }
}

// reverse samples: the S nearest pushers of each kind, appended when not already present
__global__ __launch_bounds__(64) void k_rsample(KnnArgs a) {
  __shared__ uint64_t buf[256];
  __shared__ uint32_t own[64];
  const uint32_t v = blockIdx.x;
  if (v >= a.N) return;
  const int lane = threadIdx.x;
  uint32_t* h = a.H + size_t(v) * HROW;
  for (int kind = 0; kind < 2; ++kind) {
    const int c0 = kind == 0 ? a.nnew[v] : a.nold[v];
    uint32_t* list = h + (kind == 0 ? 0 : 64);
    if (lane < c0) own[lane] = list[lane];
    __syncthreads();
    const int r = pb_top(kind == 0 ? a.pn : a.po, v, S, false, [](uint64_t) { return true; }, buf);
    bool add = false;
    uint32_t w = 0;
    if (lane < r) {
      w = key_id(buf[lane]);
      add = true;
      for (int j = 0; j < c0; ++j) add = add && (own[j] != w);
    }
    const uint64_t m = ballot(add);
    if (add) list[c0 + mask_rank(m)] = w;
    if (lane == 0) {
      if (kind == 0) a.na[v] = c0 + __popcll(m);
      else a.nb[v] = c0 + __popcll(m);
    }
    __syncthreads();
  }
}

// local join, exact canonical distances; one 256-thread workgroup (16 groups) per node
template <int MC>
__global__ __launch_bounds__(256) void k_localjoin_exact(KnnArgs a) {
  __shared__ uint32_t A[64], B[64];
  __shared__ float wA[64], wB[64];
  __shared__ float MAB[64 * 64], MAA[64 * 64];
  const uint32_t u = row_of_block(a.perm, a.N);
  if (u >= a.N) return;
  const int tid = threadIdx.x, gg = tid >> 4, t = tid & 15;
  const int K = a.K, D = a.D;
  const int nA = a.na[u], nB = a.nb[u];
  if (tid < 64) {
    if (tid < nA) {
      A[tid] = a.H[size_t(u) * HROW + tid];
      wA[tid] = a.GD[size_t(A[tid]) * K + K - 1];
    }
    if (tid < nB) {
      B[tid] = a.H[size_t(u) * HROW + 64 + tid];
      wB[tid] = a.GD[size_t(B[tid]) * K + K - 1];
    }
  }
  __syncthreads();
  const float INF = __uint_as_float(0x7F800000u);
  for (int i = gg; i < nA; i += 16) {
    Row16<MC> ra;
    load_row16<MC>(ra, a.X + size_t(A[i]) * D, D, t);
    // row i against B[0..nB) then A[i+1..nA), four rows in flight per group
    const int nj = nB + (nA - i - 1);
    for (int j0 = 0; j0 < nj; j0 += 4) {
      const float* rows[4];
      bool same[4];
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int j = j0 + b;
        rows[b] = a.X;  // unused slots read row 0
        same[b] = false;
        if (j < nB) {
          same[b] = A[i] == B[j];
          rows[b] = a.X + size_t(B[j]) * D;
        } else if (j < nj) {
          rows[b] = a.X + size_t(A[i + 1 + (j - nB)]) * D;
        }
      }
      float dd[4];
      dist16_batch<MC, 4>(ra, rows, D, t, dd);
      if (t == 0) {
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const int j = j0 + b;
          if (j < nB) {
            MAB[i * 64 + j] = same[b] ? INF : dd[b];
          } else if (j < nj) {
            const int jj = i + 1 + (j - nB);
            MAA[i * 64 + jj] = dd[b];
            MAA[jj * 64 + i] = dd[b];
          }
        }
      }
    }
    if (t == 0) MAA[i * 64 + i] = INF;
  }
  __syncthreads();
  const int r = tid >> 2, tile = tid & 3, lo = tile * TILE;
  // rows of A against tiles of B
  if (r < nA && lo < nB) {
    uint64_t best = TG_KEY_INF;
    for (int j = lo; j < min(nB, lo + TILE); ++j) {
      const float d = MAB[r * 64 + j];
      if (d < INF) best = min(best, mk_key(d, B[j]));
    }
    if (best != TG_KEY_INF && key_d(best) < wA[r] && !row_has_key(a.G, a.GD, K, A[r], best))
      pb_push(a.pc, A[r], best);
  }
  // columns of B against tiles of A
  if (r < nB && lo < nA) {
    uint64_t best = TG_KEY_INF;
    for (int i = lo; i < min(nA, lo + TILE); ++i) {
      const float d = MAB[i * 64 + r];
      if (d < INF) best = min(best, mk_key(d, A[i]));
    }
    if (best != TG_KEY_INF && key_d(best) < wB[r] && !row_has_key(a.G, a.GD, K, B[r], best))
      pb_push(a.pc, B[r], best);
  }
  // rows of A against tiles of A
  if (r < nA && lo < nA) {
    uint64_t best = TG_KEY_INF;
    for (int j = lo; j < min(nA, lo + TILE); ++j) {
      const float d = MAA[r * 64 + j];
      if (d < INF) best = min(best, mk_key(d, A[j]));
    }
    if (best != TG_KEY_INF && key_d(best) < wA[r] && !row_has_key(a.G, a.GD, K, A[r], best))
      pb_push(a.pc, A[r], best);
  }
  if (a.counters && tid == 0)
    atomicAdd(&cnt_slot(a.counters)[0], (unsigned long long)(nA * nB + nA * (nA - 1) / 2));
}

// ------------------------------------------------------------------ row merge helper
// Keeps the K smallest of the row (keys gk[0..K), flags vf) and the candidates cand[0..nc)
// (sorted, unique, none already listed); writes ids/dists/flags to the output row.
__device__ void merge_row(int K, const uint64_t* gk, const uint8_t* vf, const uint64_t* cand,
                          int nc, uint32_t* oid, float* od, uint8_t* ov) {
  const int lane = threadIdx.x;
  for (int i = lane; i < K; i += TG_WAVE) {
    const uint64_t k = gk[i];
    const int pos = i + lower_bound_lds(cand, nc, k);
    if (pos < K) {
      oid[pos] = key_id(k);
      od[pos] = key_d(k);
      ov[pos] = vf[i];
    }
  }
  for (int b = lane; b < nc; b += TG_WAVE) {
    const uint64_t k = cand[b];
    const int pos = b + lower_bound_lds(gk, K, k);
    if (pos < K) {
      oid[pos] = key_id(k);
      od[pos] = key_d(k);
      ov[pos] = 0;
    }
  }
}

struct RowLds {
  uint64_t gk[128];
  uint8_t vf[128];
};

// loads row u into LDS: its keys (sorted by (d, id), as every list is) and visit flags
__device__ void load_row_lds(const KnnArgs& a, const uint32_t* G, const float* GD, uint32_t u,
                             RowLds& r) {
  const int lane = threadIdx.x, K = a.K;
  for (int j = lane; j < K; j += TG_WAVE) {
    r.gk[j] = mk_key(GD[size_t(u) * K + j], G[size_t(u) * K + j]);
    r.vf[j] = a.V[size_t(u) * K + j];
  }
  __syncthreads();
}

// merge of the local-join candidates
__global__ __launch_bounds__(64) void k_merge(KnnArgs a) {
  __shared__ RowLds row;
  __shared__ uint64_t buf[256];
  const uint32_t v = blockIdx.x;
  if (v >= a.N) return;
  const int K = a.K;
  // Every push was checked by the local join against this very row (the merge runs after it,
  // before any row changes) and entered only below its worst distance and when not listed;
  // a build's distances are canonical (its lists start from the initial sort), so a listed
  // id would have arrived with its listed key.  No id search is needed here.
  load_row_lds(a, a.G, a.GD, v, row);
  const float worst = key_d(row.gk[K - 1]);
  const int nc = pb_top(a.pc, v, K, true,
                        [&](uint64_t k) { return key_id(k) != v && key_d(k) < worst; }, buf);
  if (nc > 0)
    merge_row(K, row.gk, row.vf, buf, nc, a.G + size_t(v) * K, a.GD + size_t(v) * K,
              a.V + size_t(v) * K);
}

// ---------------------------------------------------------------- reverse build
// The reverse lists of do_reverse_graph (Tagore_src.cu:1187-1206) as exact runs per
// destination, built without a global atomic per edge (in-degrees of a kNN graph are heavy
// tailed: a fixed number of slots per node overflows for a large share of the edges):
//   1. k_rev_hist: each tile of edges counts its destinations per block of 2^lgR ids (LDS);
//   2. an exclusive scan of those counts, block-major, gives every (block, tile) pair its
//      output range;
//   3. k_rev_part: each tile scatters its edges -- key (d, u) and destination v -- into its
//      ranges, REV_CH edges per pass: a pass is grouped by block in LDS first, so each
//      block's edges leave as one contiguous run of stores (whole cache lines, not scattered
//      8- and 4-byte writes);
//   4. k_rev_runs: per block, LDS counts and a scan give every destination its run
//      [off[v], off[v] + cnt[v]); LDS cursors place the keys.
// The runs land in the reverse PushBuf as its grouped overflow with cap 0 (cnt, ov_off,
// ov_sorted), the form pb_top reads; partitioned keys and destinations use its slot area and
// ov_dest as scratch.  The order inside a run is unspecified: pb_top selects by (d, id).
constexpr int REV_THREADS = 1024;  // 16 waves per workgroup: one workgroup per CU and tile
                                   // (hist, part) or block (runs)
constexpr uint32_t REV_MAXP = 8192;   // destination blocks (LDS counters of hist and part)
constexpr uint32_t REV_MAXR = 8192;   // ids per block (LDS counters of runs)
constexpr uint32_t REV_TILES = 256;

struct RevPlan {
  uint32_t lgR, P, tiles;
  uint64_t n, T;       // edges, edges per tile
  uint32_t* H;         // [P * tiles + 1]
};

RevPlan rev_plan(uint32_t N, uint32_t K, uint32_t* H) {
  RevPlan r;
  r.lgR = 10;  // at least 1024 ids per block, so every thread of k_rev_runs owns one
  while (r.lgR < 13 && (size_t(N) >> r.lgR) > 256) ++r.lgR;
  r.P = uint32_t((size_t(N) + (size_t(1) << r.lgR) - 1) >> r.lgR);
  r.n = uint64_t(N) * K;
  const uint64_t t = std::max<uint64_t>(1, std::min<uint64_t>(REV_TILES, (r.n + 4095) / 4096));
  r.T = (r.n + t - 1) / t;
  r.tiles = uint32_t((r.n + r.T - 1) / r.T);
  r.H = H;
  return r;
}

__global__ __launch_bounds__(REV_THREADS) void k_rev_hist(KnnArgs a, RevPlan r) {
  __shared__ uint32_t h[REV_MAXP];
  const uint32_t t = blockIdx.x;
  for (uint32_t b = threadIdx.x; b < r.P; b += REV_THREADS) h[b] = 0;
  __syncthreads();
  const uint64_t e0 = uint64_t(t) * r.T, e1 = min(r.n, e0 + r.T);
  constexpr int U = 8;  // loads per thread issued before any is used
  for (uint64_t c0 = e0; c0 < e1; c0 += uint64_t(REV_THREADS) * U) {
    uint32_t v[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const uint64_t e = c0 + uint64_t(k) * REV_THREADS + threadIdx.x;
      v[k] = e < e1 ? a.G[e] : TG_INV;
    }
#pragma unroll
    for (int k = 0; k < U; ++k)
      if (v[k] < a.N) atomicAdd(&h[v[k] >> r.lgR], 1u);
  }
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < r.P; b += REV_THREADS) r.H[size_t(b) * r.tiles + t] = h[b];
}

constexpr int REV_CH = 4096;  // edges per pass of k_rev_part
static_assert(REV_CH % REV_THREADS == 0, "a pass is a whole number of edges per thread");

// Exclusive scan in place of x[0..n) (n <= REV_MAXP) by a REV_THREADS-thread workgroup, each
// thread owning a run of consecutive entries; returns the total (in every thread).
__device__ uint32_t rev_block_scan(uint32_t* x, uint32_t n, uint32_t* wsum) {
  const uint32_t per = (n + REV_THREADS - 1) / REV_THREADS, i0 = threadIdx.x * per;
  uint32_t own = 0;
  for (uint32_t i = 0; i < per; ++i)
    if (i0 + i < n) own += x[i0 + i];
  const uint32_t incl = uint32_t(wave_incl_scan(int(own)));
  const int w = threadIdx.x / TG_WAVE;
  if (lane_id() == TG_WAVE - 1) wsum[w] = incl;
  __syncthreads();
  uint32_t run = incl - own, total = 0;
  for (int q = 0; q < REV_THREADS / TG_WAVE; ++q) {
    if (q < w) run += wsum[q];
    total += wsum[q];
  }
  for (uint32_t i = 0; i < per; ++i)
    if (i0 + i < n) {
      const uint32_t c = x[i0 + i];
      x[i0 + i] = run;
      run += c;
    }
  __syncthreads();
  return total;
}

__global__ __launch_bounds__(REV_THREADS) void k_rev_part(KnnArgs a, RevPlan r) {
  __shared__ uint32_t cur[REV_MAXP];  // this tile's next output position of every block
  __shared__ uint32_t off[REV_MAXP];  // a pass's counts per block, then their offsets
  __shared__ uint64_t skey[REV_CH];   // a pass's keys and destinations, grouped by block
  __shared__ uint32_t sv[REV_CH];
  __shared__ uint32_t wsum[REV_THREADS / TG_WAVE];
  constexpr int EPT = REV_CH / REV_THREADS;
  const uint32_t t = blockIdx.x, tid = threadIdx.x;
  for (uint32_t b = tid; b < r.P; b += REV_THREADS) cur[b] = r.H[size_t(b) * r.tiles + t];
  uint64_t* pkey = a.pr.slots;
  uint32_t* pv = a.pr.ov_dest;
  const uint64_t e0 = uint64_t(t) * r.T, e1 = min(r.n, e0 + r.T);
  for (uint64_t c0 = e0; c0 < e1; c0 += REV_CH) {
    for (uint32_t b = tid; b < r.P; b += REV_THREADS) off[b] = 0;
    __syncthreads();
    uint32_t v[EPT], rk[EPT];
    uint64_t key[EPT];
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const uint64_t e = c0 + uint64_t(k) * REV_THREADS + tid;
      v[k] = e < e1 ? a.G[e] : TG_INV;
      key[k] = 0;
      rk[k] = 0;
      if (v[k] < a.N) {
        key[k] = mk_key(a.GD[e], uint32_t(e / a.K));
        rk[k] = atomicAdd(&off[v[k] >> r.lgR], 1u);
      }
    }
    __syncthreads();
    const uint32_t m = rev_block_scan(off, r.P, wsum);
#pragma unroll
    for (int k = 0; k < EPT; ++k)
      if (v[k] < a.N) {
        const uint32_t p = off[v[k] >> r.lgR] + rk[k];
        skey[p] = key[k];
        sv[p] = v[k];
      }
    __syncthreads();
    // each block's run of the pass to consecutive positions from its cursor
    for (uint32_t i = tid; i < m; i += REV_THREADS) {
      const uint32_t vv = sv[i], b = vv >> r.lgR;
      const uint32_t p = cur[b] + (i - off[b]);
      if (p < r.n) {  // always, for the counts of k_rev_hist
        pkey[p] = skey[i];
        pv[p] = vv;
      }
    }
    __syncthreads();
    for (uint32_t b = tid; b < r.P; b += REV_THREADS)
      cur[b] += (b + 1 < r.P ? off[b + 1] : m) - off[b];
    __syncthreads();
  }
}

__global__ __launch_bounds__(REV_THREADS) void k_rev_runs(KnnArgs a, RevPlan r) {
  __shared__ uint32_t c[REV_MAXR];
  __shared__ uint32_t wsum[REV_THREADS / TG_WAVE];
  const uint32_t b = blockIdx.x, R = 1u << r.lgR, v0 = b << r.lgR;
  const uint32_t end = r.H[size_t(r.P) * r.tiles];
  uint32_t hi = b + 1 < r.P ? r.H[size_t(b + 1) * r.tiles] : end;
  uint32_t lo = r.H[size_t(b) * r.tiles];
  hi = min(hi, uint32_t(r.n));
  lo = min(lo, hi);
  const uint32_t* pv = a.pr.ov_dest;
  const uint64_t* pkey = a.pr.slots;
  // RUN_U entries per thread and step: their loads are all issued before any is used
  constexpr uint32_t RUN_U = 8;
  for (uint32_t i = threadIdx.x; i < R; i += REV_THREADS) c[i] = 0;
  __syncthreads();
  for (uint64_t p0 = lo; p0 < hi; p0 += REV_THREADS * RUN_U) {
    uint32_t l[RUN_U];
#pragma unroll
    for (uint32_t k = 0; k < RUN_U; ++k) {
      const uint64_t p = p0 + k * REV_THREADS + threadIdx.x;
      l[k] = p < hi ? pv[p] - v0 : R;
    }
#pragma unroll
    for (uint32_t k = 0; k < RUN_U; ++k)
      if (l[k] < R) atomicAdd(&c[l[k]], 1u);
  }
  __syncthreads();
  // exclusive scan of c[0..R): thread i owns c[i*per .. (i+1)*per)
  const uint32_t per = R / REV_THREADS, i0 = threadIdx.x * per;
  uint32_t own = 0;
  for (uint32_t i = 0; i < per; ++i) own += c[i0 + i];
  const uint32_t incl = uint32_t(wave_incl_scan(int(own)));
  const int w = threadIdx.x / TG_WAVE;
  if (lane_id() == TG_WAVE - 1) wsum[w] = incl;
  __syncthreads();
  uint32_t run = incl - own;
  for (int q = 0; q < w; ++q) run += wsum[q];
  for (uint32_t i = 0; i < per; ++i) {
    const uint32_t ci = c[i0 + i], v = v0 + i0 + i;
    if (v < a.N) {
      a.pr.cnt[v] = ci;
      a.pr.ov_off[v] = lo + run;
    }
    c[i0 + i] = run;
    run += ci;
  }
  if (b + 1 == r.P && threadIdx.x == 0) a.pr.ov_off[a.N] = hi;
  __syncthreads();
  for (uint64_t p0 = lo; p0 < hi; p0 += REV_THREADS * RUN_U) {
    uint32_t l[RUN_U];
    uint64_t key[RUN_U];
#pragma unroll
    for (uint32_t k = 0; k < RUN_U; ++k) {
      const uint64_t p = p0 + k * REV_THREADS + threadIdx.x;
      l[k] = p < hi ? pv[p] - v0 : R;
      key[k] = p < hi ? pkey[p] : 0ull;
    }
#pragma unroll
    for (uint32_t k = 0; k < RUN_U; ++k)
      if (l[k] < R) {
        const uint32_t q = atomicAdd(&c[l[k]], 1u);
        if (lo + q < hi) a.pr.ov_sorted[lo + q] = key[k];
      }
  }
}

__global__ __launch_bounds__(64) void k_rev_reduce(KnnArgs a) {
  __shared__ uint64_t buf[256];
  const uint32_t v = blockIdx.x;
  if (v >= a.N) return;
  const int K = a.K;
  const int n = pb_top(a.pr, v, K, true, [](uint64_t) { return true; }, buf);
  for (int i = threadIdx.x; i < K; i += TG_WAVE) {
    a.R0[size_t(v) * K + i] = i < n ? buf[i] : TG_KEY_INF;
    a.V[size_t(v) * K + i] = 0;
  }
  if (threadIdx.x == 0) a.Rn0[v] = n;
}

// ------------------------------------------------------------------------- phase 2
// BF16 screen of one candidate: lane t of the 16-lane group holds u' elements 128m + 8t..+7.
template <int MB>
struct RowBf {
  float v[MB][8];
};
__device__ __forceinline__ float bfbits2f(uint32_t bits16) { return __uint_as_float(bits16 << 16); }

template <int MB>
__device__ __forceinline__ void load_rowbf(RowBf<MB>& r, const uint16_t* x, int D, int t) {
#pragma unroll
  for (int m = 0; m < MB; ++m) {
    const int i = 128 * m + 8 * t;
    uint4 q = make_uint4(0, 0, 0, 0);
    if (i < D) q = *reinterpret_cast<const uint4*>(x + i);
    const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      r.v[m][2 * k] = bfbits2f(w[k] & 0xFFFFu);
      r.v[m][2 * k + 1] = bfbits2f(w[k] >> 16);
    }
  }
}

template <int MB>
__device__ __forceinline__ float dot_bf16(const RowBf<MB>& a, const uint16_t* y, int D, int t) {
  float s = 0.f;
#pragma unroll
  for (int m = 0; m < MB; ++m) {
    const int i = 128 * m + 8 * t;
    if (i < D) {
      const uint4 q = *reinterpret_cast<const uint4*>(y + i);
      const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        s = __builtin_fmaf(a.v[m][2 * k], bfbits2f(w[k] & 0xFFFFu), s);
        s = __builtin_fmaf(a.v[m][2 * k + 1], bfbits2f(w[k] >> 16), s);
      }
    }
  }
#pragma unroll
  for (int o = 8; o >= 1; o >>= 1) s += __shfl_xor(s, o, 16);
  return s;
}

// FP8 screen of one candidate: lane t of the 16-lane group holds u's dequantised elements
// 256m + 16t..+15 (D % 16 == 0); the scales are applied by the caller.
template <int MF>
struct RowF8 {
  float v[MF][16];
};
typedef float f32x2 __attribute__((ext_vector_type(2)));

template <int MF>
__device__ __forceinline__ void load_rowf8(RowF8<MF>& r, const uint8_t* x, int D, int t) {
#pragma unroll
  for (int m = 0; m < MF; ++m) {
    const int i = 256 * m + 16 * t;
    uint4 q = make_uint4(0, 0, 0, 0);
    if (i < D) q = *reinterpret_cast<const uint4*>(x + i);
    const int w[4] = {int(q.x), int(q.y), int(q.z), int(q.w)};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const f32x2 lo = __builtin_amdgcn_cvt_pk_f32_fp8(w[k], false);
      const f32x2 hi = __builtin_amdgcn_cvt_pk_f32_fp8(w[k], true);
      r.v[m][4 * k] = lo.x;
      r.v[m][4 * k + 1] = lo.y;
      r.v[m][4 * k + 2] = hi.x;
      r.v[m][4 * k + 3] = hi.y;
    }
  }
}

template <int MF>
__device__ __forceinline__ float dot_f8(const RowF8<MF>& a, const uint8_t* y, int D, int t) {
  float s = 0.f;
#pragma unroll
  for (int m = 0; m < MF; ++m) {
    const int i = 256 * m + 16 * t;
    if (i < D) {
      const uint4 q = *reinterpret_cast<const uint4*>(y + i);
      const int w[4] = {int(q.x), int(q.y), int(q.z), int(q.w)};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const f32x2 lo = __builtin_amdgcn_cvt_pk_f32_fp8(w[k], false);
        const f32x2 hi = __builtin_amdgcn_cvt_pk_f32_fp8(w[k], true);
        s = __builtin_fmaf(a.v[m][4 * k], lo.x, s);
        s = __builtin_fmaf(a.v[m][4 * k + 1], lo.y, s);
        s = __builtin_fmaf(a.v[m][4 * k + 2], hi.x, s);
        s = __builtin_fmaf(a.v[m][4 * k + 3], hi.y, s);
      }
    }
  }
#pragma unroll
  for (int o = 8; o >= 1; o >>= 1) s += __shfl_xor(s, o, 16);
  return s;
}

// Four candidates at a time (their loads issued before any arithmetic); each result is
// the same operation sequence as dot_bf16 / dot_f8 of that candidate.
template <int MB>
__device__ __forceinline__ void dot_bf16x4(const RowBf<MB>& a, const uint16_t* const (&y)[4],
                                           int D, int t, float (&out)[4]) {
  uint4 q[4][MB];
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int m = 0; m < MB; ++m) {
      const int i = 128 * m + 8 * t;
      q[b][m] = i < D ? *reinterpret_cast<const uint4*>(y[b] + i) : make_uint4(0, 0, 0, 0);
    }
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    float s = 0.f;
#pragma unroll
    for (int m = 0; m < MB; ++m) {
      if (128 * m + 8 * t < D) {
        const uint32_t w[4] = {q[b][m].x, q[b][m].y, q[b][m].z, q[b][m].w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          s = __builtin_fmaf(a.v[m][2 * k], bfbits2f(w[k] & 0xFFFFu), s);
          s = __builtin_fmaf(a.v[m][2 * k + 1], bfbits2f(w[k] >> 16), s);
        }
      }
    }
    out[b] = s;
  }
#pragma unroll
  for (int o = 8; o >= 1; o >>= 1)
#pragma unroll
    for (int b = 0; b < 4; ++b) out[b] += __shfl_xor(out[b], o, 16);
}

template <int MF>
__device__ __forceinline__ void dot_f8x4(const RowF8<MF>& a, const uint8_t* const (&y)[4], int D,
                                         int t, float (&out)[4]) {
  uint4 q[4][MF];
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int m = 0; m < MF; ++m) {
      const int i = 256 * m + 16 * t;
      q[b][m] = i < D ? *reinterpret_cast<const uint4*>(y[b] + i) : make_uint4(0, 0, 0, 0);
    }
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    float s = 0.f;
#pragma unroll
    for (int m = 0; m < MF; ++m) {
      if (256 * m + 16 * t < D) {
        const int w[4] = {int(q[b][m].x), int(q[b][m].y), int(q[b][m].z), int(q[b][m].w)};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const f32x2 lo = __builtin_amdgcn_cvt_pk_f32_fp8(w[k], false);
          const f32x2 hi = __builtin_amdgcn_cvt_pk_f32_fp8(w[k], true);
          s = __builtin_fmaf(a.v[m][4 * k], lo.x, s);
          s = __builtin_fmaf(a.v[m][4 * k + 1], lo.y, s);
          s = __builtin_fmaf(a.v[m][4 * k + 2], hi.x, s);
          s = __builtin_fmaf(a.v[m][4 * k + 3], hi.y, s);
        }
      }
    }
    out[b] = s;
  }
#pragma unroll
  for (int o = 8; o >= 1; o >>= 1)
#pragma unroll
    for (int b = 0; b < 4; ++b) out[b] += __shfl_xor(out[b], o, 16);
}

__device__ __forceinline__ float screen_bound2(float na, float qa, float ea, float nb, float qb,
                                               float eb, float eta) {
  return 2.0f * (ea * qb + qa * eb + ea * eb) + eta * (na + nb + 2.0f * qa * qb);
}

template <int MC, int PREC, class T>
__global__ __launch_bounds__(64) void k_phase2(KnnArgs a) {
  __shared__ RowLds row;
  __shared__ uint32_t ids[512];  // gathered ids, then the candidates compacted in place
  __shared__ uint64_t kv[512];
  __shared__ uint32_t lead[LEADERS];
  const uint32_t u = row_of_block(a.perm, a.N);
  if (u >= a.N) return;
  const int lane = threadIdx.x, g = lane >> 4, t = lane & 15;
  const int K = a.K, D = a.D;
  load_row_lds(a, a.G, a.GD, u, row);
  // leaders: first LEADERS unvisited positions in list order
  int nl = 0;
  for (int base = 0; base < K && nl < LEADERS; base += TG_WAVE) {
    const int j = base + lane;
    const bool unv = j < K && row.vf[j] == 0;
    const uint64_t m = ballot(unv);
    const int rk = nl + mask_rank(m);
    if (unv && rk < LEADERS) {
      lead[rk] = key_id(row.gk[j]);
      row.vf[j] = 1;
    }
    nl = min(LEADERS, nl + int(__popcll(m)));
  }
  __syncthreads();
  // gather candidate ids: the leaders' counts first, then every leader's loads back to back
  const int nf = min(TAKE, K);
  const int cnt_l = lane < nl ? nf + min(TAKE, int(a.Rn0[lead[lane]])) : 0;
  int off[LEADERS + 1];
  off[0] = 0;
#pragma unroll
  for (int l = 0; l < LEADERS; ++l) off[l + 1] = off[l] + __shfl(cnt_l, l, TG_WAVE);
  const int tot = off[LEADERS];
#pragma unroll
  for (int l = 0; l < LEADERS; ++l) {
    if (l >= nl) break;
    const uint32_t v = lead[l];
    for (int i = lane; i < off[l + 1] - off[l]; i += TG_WAVE)
      ids[off[l] + i] =
          i < nf ? a.G[size_t(v) * K + i] : key_id(a.R0[size_t(v) * K + (i - nf)]);
  }
  // unique, not u, not listed, in gather order (the keys are sorted afterwards, so the order
  // of evaluation does not matter): an LDS id set in kv's space, 1024 slots for the K listed
  // ids, inserted first, and at most 512 gathered ones; the insert that finds the slot empty
  // keeps the id
  uint32_t* hs = reinterpret_cast<uint32_t*>(kv);
  for (int i = lane; i < 1024; i += TG_WAVE) hs[i] = TG_INV;
  __syncthreads();
  for (int j = lane; j < K; j += TG_WAVE) {
    const uint32_t id = key_id(row.gk[j]);
    uint32_t h = (id * 0x9E3779B1u) >> 22;
    for (;;) {
      const uint32_t old = atomicCAS(&hs[h], TG_INV, id);
      if (old == TG_INV || old == id) break;
      h = (h + 1) & 1023u;
    }
  }
  __syncthreads();
  int cnt = 0;
  for (int base = 0; base < tot; base += TG_WAVE) {
    const int i = base + lane;
    bool ok = false;
    const uint32_t id = i < tot ? ids[i] : TG_INV;
    {
      if (id != TG_INV && id != u) {
        uint32_t h = (id * 0x9E3779B1u) >> 22;
        for (;;) {
          const uint32_t old = atomicCAS(&hs[h], TG_INV, id);
          if (old == TG_INV) {
            ok = true;
            break;
          }
          if (old == id) break;
          h = (h + 1) & 1023u;
        }
      }
    }
    // in place: this chunk was read above, and writes land at or below the read positions
    const uint64_t m = ballot(ok);
    if (ok) ids[cnt + mask_rank(m)] = id;
    cnt += __popcll(m);
  }
  __syncthreads();
  Row16<MC> xu;
  load_row16<MC>(xu, a.X + size_t(u) * D, D, t);
  const float worst = key_d(row.gk[K - 1]);
  const int nk = pow2_at_least(max(cnt, 1));
  constexpr int MB = (MC + 1) / 2;
  constexpr int MF = (MC + 3) / 4;
  constexpr bool SCREEN = PREC == KP_BF16 || PREC == KP_FP8;
  RowBf<PREC == KP_BF16 ? MB : 1> ub;
  RowF8<PREC == KP_FP8 ? MF : 1> uf;
  float nu = 0.f, qu = 0.f, eu = 0.f, su = 1.f;
  if (PREC == KP_BF16) load_rowbf(ub, a.Xb + size_t(u) * D, D, t);
  if (PREC == KP_FP8) {
    load_rowf8(uf, a.Xq8 + size_t(u) * D, D, t);
    su = a.sc8[u];
  }
  if (SCREEN) {
    nu = a.nrm2[u];
    qu = a.nq[u];
    eu = a.ne[u];
  }
  for (int q = lane; q < nk; q += TG_WAVE) kv[q] = TG_KEY_INF;
  // BF16 screen: keep (in order) only the candidates whose bound interval reaches below the
  // worst distance; everything kept is re-scored exactly.  The keep flags use kv's space
  // (the keys are cleared again after the compaction).
  int nre = cnt;
  bool done = false;
  if constexpr (PREC == KP_BF16 && MC <= 4) {
    if (a.exact_screen && a.X8) {
      // integer data in [0, 255] (D <= 256): exact distances from the u8 copy (quad layout,
      // 32 candidates per step), the canonical distances themselves
      Row16<MC> xu8;
      load_row16<MC>(xu8, a.X8 + size_t(u) * D, D, t);
      __syncthreads();
      eval_batch<MC>(a.X8, D, xu8, cnt, [&](int q) { return ids[q]; },
                     [&](int q, float d) {
                       if (d < worst) kv[q] = mk_key(d, ids[q]);
                     });
      nre = 0;
      done = true;
    }
  }
  if (SCREEN && !done) {
    uint16_t* keep = reinterpret_cast<uint16_t*>(kv);
    // item q0 + g + 4b in group g, four candidates per group in flight
    for (int q0 = 0; q0 < cnt; q0 += 16) {
      uint32_t y[4];
      float dp[4], nb[4];
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int q = q0 + g + 4 * b;
        y[b] = q < cnt ? ids[q] : u;  // unused slots read row u
        nb[b] = a.nrm2[y[b]];
      }
      if (PREC == KP_FP8) {
        const uint8_t* r8[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) r8[b] = a.Xq8 + size_t(y[b]) * D;
        dot_f8x4(uf, r8, D, t, dp);
#pragma unroll
        for (int b = 0; b < 4; ++b) dp[b] = (su * a.sc8[y[b]]) * dp[b];
      } else {
        const uint16_t* rb[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) rb[b] = a.Xb + size_t(y[b]) * D;
        dot_bf16x4(ub, rb, D, t, dp);
      }
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int q = q0 + g + 4 * b;
        if (q >= cnt) continue;
        const float s = (nu + nb[b]) - 2.0f * dp[b];
        if (a.exact_screen) {
          // integer data: s is the canonical distance itself
          if (t == 0 && s < worst) kv[q] = mk_key(s, y[b]);
        } else {
          const float E = screen_bound2(nu, qu, eu, nb[b], a.nq[y[b]], a.ne[y[b]], a.eta);
          if (t == 0) keep[q] = (s - E < worst) ? 1 : 0;
        }
      }
    }
    __syncthreads();
    nre = 0;
    if (!a.exact_screen) {
      for (int base = 0; base < cnt; base += TG_WAVE) {
        const int q = base + lane;
        const bool k = q < cnt && keep[q];
        const uint32_t c = q < cnt ? ids[q] : 0u;
        const uint64_t m = ballot(k);
        __syncthreads();
        if (k) ids[nre + mask_rank(m)] = c;
        nre += __popcll(m);
        __syncthreads();
      }
      for (int q = lane; q < nk; q += TG_WAVE) kv[q] = TG_KEY_INF;
    }
  }
  __syncthreads();
  const T* Xr = a.Xr16 ? reinterpret_cast<const T*>(a.Xr16) : reinterpret_cast<const T*>(a.X);
  eval_batch<MC>(Xr, D, xu, nre, [&](int q) { return ids[q]; },
                 [&](int q, float d) {
                   if (d < worst) kv[q] = mk_key(d, ids[q]);
                 });
  if (a.counters && lane == 0 && nre) atomicAdd(&cnt_slot(a.counters)[1], (unsigned long long)nre);
  __syncthreads();
  // the keys that beat the worst entry (usually few) compacted to the front, in place (a
  // chunk is read before any of its positions is written), then sorted
  int nv = 0;
  for (int base = 0; base < nk; base += TG_WAVE) {
    const int i = base + lane;
    const uint64_t k = i < nk ? kv[i] : TG_KEY_INF;
    const bool ok = k != TG_KEY_INF;
    const uint64_t m = ballot(ok);
    if (ok) kv[nv + mask_rank(m)] = k;
    nv += int(__popcll(m));
  }
  const int ns = pow2_at_least(max(nv, 1));
  for (int i = nv + lane; i < ns; i += TG_WAVE) kv[i] = TG_KEY_INF;
  __syncthreads();
  wave_sort<512>(kv, ns);
  const int nc = min(nv, K);
  merge_row(K, row.gk, row.vf, kv, nc, a.G1 + size_t(u) * K, a.GD1 + size_t(u) * K,
            a.V + size_t(u) * K);
  // reverse pushes for entries that entered the list
  for (int b = lane; b < nc; b += TG_WAVE) {
    const uint64_t k = kv[b];
    const int pos = b + lower_bound_lds(row.gk, K, k);
    if (pos < K) pb_push(a.pr, key_id(k), mk_key(key_d(k), u));
  }
  if (a.counters && lane == 0) atomicAdd(&cnt_slot(a.counters)[0], (unsigned long long)cnt);
}

// R1[w] = K smallest of R0[w] and this iteration's pushes
__global__ __launch_bounds__(64) void k_rupdate(KnnArgs a) {
  __shared__ uint64_t buf[256];
  __shared__ uint64_t seed[128];
  const uint32_t w = blockIdx.x;
  if (w >= a.N) return;
  const int K = a.K;
  const int n0 = a.Rn0[w];
  for (int i = threadIdx.x; i < n0; i += TG_WAVE) seed[i] = a.R0[size_t(w) * K + i];
  __syncthreads();
  const int n = pb_top(a.pr, w, K, true, [](uint64_t) { return true; }, buf, seed, n0);
  for (int i = threadIdx.x; i < K; i += TG_WAVE)
    a.R1[size_t(w) * K + i] = i < n ? buf[i] : TG_KEY_INF;
  if (threadIdx.x == 0) a.Rn1[w] = n;
}

// final merge of the reverse lists into the forward lists
__global__ __launch_bounds__(64) void k_final(KnnArgs a) {
  __shared__ RowLds row;
  __shared__ uint64_t cand[128];
  const uint32_t u = blockIdx.x;
  if (u >= a.N) return;
  const int lane = threadIdx.x, K = a.K;
  // A reverse entry (d(w, u), w) carries the canonical distance, which is symmetric, so w is
  // listed exactly when that key is in the row: a search of the sorted keys, no id sort.
  load_row_lds(a, a.G, a.GD, u, row);
  const float worst = key_d(row.gk[K - 1]);
  const int nr = a.Rn0[u];
  int nc = 0;
  for (int base = 0; base < nr; base += TG_WAVE) {
    const int i = base + lane;
    bool ok = false;
    uint64_t k = 0;
    if (i < nr) {
      k = a.R0[size_t(u) * K + i];
      const int p = lower_bound_lds(row.gk, K, k);
      ok = key_id(k) != u && key_d(k) < worst && !(p < K && row.gk[p] == k);
    }
    const uint64_t m = ballot(ok);
    if (ok) cand[nc + mask_rank(m)] = k;
    nc += __popcll(m);
  }
  __syncthreads();
  if (nc > 0)
    merge_row(K, row.gk, row.vf, cand, nc, a.G + size_t(u) * K, a.GD + size_t(u) * K,
              a.V + size_t(u) * K);
}

// ------------------------------------------------------------------------- launchers
#define TG_MC_DISPATCH(KERNEL, GRID, BLOCK, S, ARGS)                 \
  do {                                                               \
    const int mc_ = (int(a.D) + 63) / 64;                            \
    if (mc_ <= 2) KERNEL<2><<<GRID, BLOCK, 0, S>>> ARGS;             \
    else if (mc_ <= 4) KERNEL<4><<<GRID, BLOCK, 0, S>>> ARGS;        \
    else KERNEL<16><<<GRID, BLOCK, 0, S>>> ARGS;                     \
  } while (0)

static void launch_init_sort(const KnnArgs& a, hipStream_t s) {
  const int mc = (int(a.D) + 63) / 64;
  const uint32_t grid = grid_rows(a.perm, a.N);
#define IS(M, T) k_init_sort<M, T><<<grid, 64, 0, s>>>(a)
#define IS_MC(T)                                                     \
  do {                                                               \
    if (mc <= 2) IS(2, T); else if (mc <= 4) IS(4, T); else IS(16, T); \
  } while (0)
  if (a.X8 && a.D <= 256) {  // exact dot4 distances on the u8 copy
    if (mc <= 2) IS(2, uint8_t);
    else IS(4, uint8_t);
  } else if (a.Xr16) {
    IS_MC(uint16_t);
  } else {
    IS_MC(float);
  }
#undef IS_MC
#undef IS
}
void launch_knn_init_sort(const KnnArgs& a, hipStream_t s) { launch_init_sort(a, s); }
void launch_knn_init(const KnnArgs& a, hipStream_t s) {
  k_init_ids<<<a.N, 64, 0, s>>>(a);
  launch_init_sort(a, s);
}
void launch_knn_sample(const KnnArgs& a, hipStream_t s) { k_sample<<<a.N, 64, 0, s>>>(a); }
void launch_knn_rsample(const KnnArgs& a, hipStream_t s) { k_rsample<<<a.N, 64, 0, s>>>(a); }
void launch_knn_localjoin_screen(const KnnArgs& a, int precision, int tile, hipStream_t s);

void launch_knn_localjoin(const KnnArgs& a, int precision, int tile, hipStream_t s) {
  if (precision == KP_EXACT)
    TG_MC_DISPATCH(k_localjoin_exact, grid_rows(a.perm, a.N), 256, s, (a));
  else
    launch_knn_localjoin_screen(a, precision, tile, s);
}
void launch_knn_merge(const KnnArgs& a, hipStream_t s) { k_merge<<<a.N, 64, 0, s>>>(a); }
size_t rev_hist_words(uint32_t N, uint32_t K) {
  const RevPlan r = rev_plan(N, K, nullptr);
  return size_t(r.P) * r.tiles + 1;
}
int launch_knn_rev_build(const KnnArgs& a, uint32_t* hist, uint32_t* scan_tmp, hipStream_t s) {
  const RevPlan r = rev_plan(a.N, a.K, hist);
  if (r.P > REV_MAXP || r.n >= a.pr.ov_cap + uint64_t(1) || r.n > uint64_t(a.N) * a.pr.cap)
    return -1;  // beyond 2^26 ids or 2^32 edges: the sizes the state allocates cannot hold it
  k_rev_hist<<<r.tiles, REV_THREADS, 0, s>>>(a, r);
  excl_scan_u32(hist, r.P * r.tiles, scan_tmp, s);
  k_rev_part<<<r.tiles, REV_THREADS, 0, s>>>(a, r);
  k_rev_runs<<<r.P, REV_THREADS, 0, s>>>(a, r);
  return 0;
}
void launch_knn_rev_reduce(const KnnArgs& a, hipStream_t s) { k_rev_reduce<<<a.N, 64, 0, s>>>(a); }
void launch_knn_phase2(const KnnArgs& a, int precision, hipStream_t s) {
  const int mc = (int(a.D) + 63) / 64;
#define P2(M, P, T) k_phase2<M, P, T><<<grid_rows(a.perm, a.N), 64, 0, s>>>(a)
#define P2_MC(P, T)                                                  \
  do {                                                               \
    if (mc <= 2) P2(2, P, T); else if (mc <= 4) P2(4, P, T); else P2(16, P, T); \
  } while (0)
  if (precision == KP_FP8) {
    if (a.Xr16) P2_MC(KP_FP8, uint16_t); else P2_MC(KP_FP8, float);
  } else if (precision == KP_BF16) {
    if (a.Xr16) P2_MC(KP_BF16, uint16_t); else P2_MC(KP_BF16, float);
  } else {
    if (a.Xr16) P2_MC(KP_EXACT, uint16_t); else P2_MC(KP_EXACT, float);
  }
#undef P2_MC
#undef P2
}
void launch_knn_rupdate(const KnnArgs& a, hipStream_t s) { k_rupdate<<<a.N, 64, 0, s>>>(a); }
void launch_knn_final(const KnnArgs& a, hipStream_t s) { k_final<<<a.N, 64, 0, s>>>(a); }

}  // namespace tg
