// search.hip -- beam search over a pruned index, the validator of the produced graphs.
//
// Restates get_neighbors (common.cu:65-127, the NSG search the reference carries but never
// runs on its outputs) for the variable-degree rows of a pruned graph, with the same rules
// as the CPU oracle's or_search:
//   * initial ids: the first min(L, deg(ep)) neighbours of the entry point, then draws from a
//     per-query SplitMix64 stream (the reference: rand() % N) that skip visited ids, until
//     min(L, N) ids or 4096 draws; all of them are marked visited;
//   * the pool holds those ids sorted by (canonical distance, id), each with a checked flag;
//   * k = 0; repeat: k = first unchecked position >= k; mark it checked and expand its node:
//     every neighbour not yet visited is marked visited and evaluated, and the pool becomes
//     the L smallest of pool and candidates -- the set the reference's one-at-a-time
//     InsertIntoPool leaves -- with nk = the position of the smallest candidate if it
//     enters (that is the smallest insertion position), else L; k = nk <= k ? nk : k + 1.
// One wave per query, the pool in registers (64*NPL >= L slots, NPL = 1, 2 or 4).  Each
// workgroup owns a visited bitmap in device memory (exact for any N) and a list of the ids
// it set; after a query it clears those words (or the whole bitmap when the list overflowed),
// so the bitmap is all zero again for the next query.  Bitmap and list are only touched with
// device-coherent atomics, loads and stores.
#include "common.hpp"
#include "kernels.hpp"

namespace tg {

namespace {

__device__ __forceinline__ uint32_t ld_dev(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dev(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// marks id visited; true when it was not before
__device__ __forceinline__ bool visit(uint32_t* bm, uint32_t id) {
  const uint32_t bit = 1u << (id & 31);
  return (atomicOr(&bm[id >> 5], bit) & bit) == 0;
}

constexpr int FILL_ROUNDS = 64;  // 64 rounds of 64 draws

template <int MC, int NPL>
__global__ __launch_bounds__(64) void k_search(SearchArgs a) {
  __shared__ uint64_t S[256];     // candidate keys
  __shared__ uint32_t nbr[256];   // initial ids, then the fresh neighbours of a round
  __shared__ uint32_t firstocc[64];
  const int lane = threadIdx.x, t = lane & 15;
  const uint32_t N = a.N, D = a.D, FD = a.FD;
  const int L = int(a.L);
  uint32_t* bm = a.bitmap + size_t(blockIdx.x) * a.words;
  uint32_t* vl = a.vlist + size_t(blockIdx.x) * a.vcap;
  const uint64_t INFKEY = mk_key(__uint_as_float(0x7F800000u), 0u);
  unsigned long long n_eval = 0, n_exp = 0;
  for (uint32_t q = blockIdx.x; q < a.nq; q += gridDim.x) {
    Row16<MC> xq;
    load_row16<MC>(xq, a.Q + size_t(q) * D, D, t);
    uint32_t nvis = 0;
    // ---- initial ids: the entry point's row, then random draws
    const int target = int(min(uint32_t(L), N));
    const int d0 = int(min(a.deg[a.ep], uint32_t(L)));
    for (int j = lane; j < d0; j += TG_WAVE) {
      const uint32_t y = a.ids[size_t(a.ep) * FD + j];
      nbr[j] = y;
      atomicOr(&bm[y >> 5], 1u << (y & 31));
      if (uint32_t(j) < a.vcap) st_dev(&vl[j], y);
    }
    int ninit = d0;
    nvis = uint32_t(d0);
    const uint64_t qs = segment_seed(a.seed, q);
    for (int r = 0; r < FILL_ROUNDS && ninit < target; ++r) {
      const uint32_t y = sm_bounded(sm_stream(qs, uint64_t(r) * 64 + lane), N);
      const bool seen = (ld_dev(&bm[y >> 5]) >> (y & 31)) & 1u;
      // first occurrence of y among this round's draws (lowest lane)
      uint64_t kv[1] = {(uint64_t(y) << 32) | uint32_t(lane)};
      reg_bitonic<1>(kv);
      const uint64_t prev = uint64_t(__shfl(int(kv[0] >> 32), (lane + 63) & 63, TG_WAVE));
      const bool first = lane == 0 || uint32_t(prev) != uint32_t(kv[0] >> 32);
      firstocc[uint32_t(kv[0])] = first ? 1u : 0u;
      __syncthreads();
      const bool ok = !seen && firstocc[lane] != 0;
      const uint64_t m = ballot(ok);
      const int need = target - ninit;
      const int rank = mask_rank(m);
      if (ok && rank < need) {
        nbr[ninit + rank] = y;
        atomicOr(&bm[y >> 5], 1u << (y & 31));
        if (nvis + uint32_t(rank) < a.vcap) st_dev(&vl[nvis + rank], y);
      }
      const int got = min(int(__popcll(m)), need);
      ninit += got;
      nvis += uint32_t(got);
      __syncthreads();
    }
    // ---- the pool: initial ids with distances, sorted
    for (int i = lane; i < ninit; i += TG_WAVE) S[i] = INFKEY | nbr[i];
    __syncthreads();
    eval_batch<MC>(a.X, int(D), xq, ninit, [&](int i) { return nbr[i]; },
                   [&](int i, float d) { S[i] = mk_key(d, nbr[i]); });
    n_eval += uint64_t(ninit);
    __syncthreads();
    uint64_t v[NPL];
    uint32_t f[NPL];
#pragma unroll
    for (int b = 0; b < NPL; ++b) {
      const int p = b * 64 + lane;
      v[b] = p < ninit ? S[p] : TG_KEY_INF;
      f[b] = 0;
    }
    reg_bitonic<NPL>(v);
    __syncthreads();
    // ---- expansions
    int kk = 0;
    for (;;) {
      int first = -1;
      uint32_t n = 0;
#pragma unroll
      for (int b = 0; b < NPL; ++b) {
        if (first >= 0) break;
        const int p = b * 64 + lane;
        const bool ok = p >= kk && p < L && v[b] != TG_KEY_INF && f[b] == 0;
        const uint64_t m = ballot(ok);
        if (m) {
          const int fl = __ffsll((unsigned long long)m) - 1;
          first = b * 64 + fl;
          n = uint32_t(__shfl(int(key_id(v[b])), fl, TG_WAVE));
          if (lane == fl) f[b] = 1;
        }
      }
      if (first < 0) break;
      ++n_exp;
      // unvisited neighbours of n, in row order
      const int dn = int(a.deg[n]);
      int cnt = 0;
#pragma unroll
      for (int base = 0; base < 128; base += TG_WAVE) {
        const int j = base + lane;
        bool fresh = false;
        uint32_t y = 0;
        if (j < dn) {
          y = a.ids[size_t(n) * FD + j];
          fresh = visit(bm, y);
        }
        const uint64_t fm = ballot(fresh);
        if (fresh) {
          const int pos = cnt + mask_rank(fm);
          nbr[pos] = y;
          if (nvis + uint32_t(pos) < a.vcap) st_dev(&vl[nvis + pos], y);
        }
        cnt += int(__popcll(fm));
      }
      nvis += uint32_t(cnt);
      __syncthreads();
      eval_batch<MC>(a.X, int(D), xq, cnt, [&](int i) { return nbr[i]; },
                     [&](int i, float d) { S[i] = mk_key(d, nbr[i]); });
      n_eval += uint64_t(cnt);
      __syncthreads();
      if (cnt == 0) {
        kk = first + 1;
        continue;
      }
      // candidates sorted in registers (at most 128)
      uint64_t c[2];
#pragma unroll
      for (int r = 0; r < 2; ++r) c[r] = r * 64 + lane < cnt ? S[r * 64 + lane] : TG_KEY_INF;
      reg_bitonic<2>(c);
      const uint64_t cmin = uint64_t(__shfl(int(c[0] >> 32), 0, TG_WAVE)) << 32 |
                            uint32_t(__shfl(int(uint32_t(c[0])), 0, TG_WAVE));
      // the pool's last slot L-1
      const int lb = (L - 1) >> 6, ll = (L - 1) & 63;
      uint64_t lastk = TG_KEY_INF;
#pragma unroll
      for (int b = 0; b < NPL; ++b)
        if (b == lb)
          lastk = uint64_t(uint32_t(__shfl(int(v[b] >> 32), ll, TG_WAVE))) << 32 |
                  uint32_t(__shfl(int(uint32_t(v[b])), ll, TG_WAVE));
      int nk = L;
      if (cmin < lastk) {
        nk = 0;
#pragma unroll
        for (int b = 0; b < NPL; ++b) nk += int(__popcll(ballot(v[b] < cmin)));
      }
      // the L smallest of pool and candidates: min(pool[p], cand[64 NPL - 1 - p]) is
      // bitonic; candidate slots past the first two are +inf
      uint64_t w[NPL];
      uint32_t g[NPL];
#pragma unroll
      for (int b = 0; b < NPL; ++b) {
        const int rs = NPL - 1 - b;
        const uint64_t rv = rs < 2 ? lane_xor64(c[rs < 2 ? rs : 0], 63) : TG_KEY_INF;
        const bool keep = v[b] < rv;
        w[b] = keep ? v[b] : rv;
        g[b] = keep ? f[b] : 0u;
      }
      reg_bitonic_merge_kv<NPL>(w, g);
#pragma unroll
      for (int b = 0; b < NPL; ++b) {
        const bool in = b * 64 + lane < L;
        v[b] = in ? w[b] : TG_KEY_INF;
        f[b] = in ? g[b] : 0u;
      }
      kk = nk <= first ? nk : first + 1;
    }
    // ---- the first k entries
#pragma unroll
    for (int b = 0; b < NPL; ++b) {
      const int p = b * 64 + lane;
      if (p < int(a.k)) {
        const bool valid = v[b] != TG_KEY_INF;
        a.out_ids[size_t(q) * a.k + p] = valid ? key_id(v[b]) : TG_INV;
        a.out_d[size_t(q) * a.k + p] = valid ? key_d(v[b]) : __uint_as_float(0x7F800000u);
      }
    }
    // ---- clear this query's bits
    if (nvis <= a.vcap) {
      for (uint32_t i = lane; i < nvis; i += TG_WAVE) st_dev(&bm[ld_dev(&vl[i]) >> 5], 0u);
    } else {
      for (uint32_t i = lane; i < a.words; i += TG_WAVE) st_dev(&bm[i], 0u);
    }
    __threadfence();
    __syncthreads();
  }
  if (a.counters && lane == 0) {
    atomicAdd(&a.counters[0], n_eval);
    atomicAdd(&a.counters[1], n_exp);
  }
}

}  // namespace

void launch_search(const SearchArgs& a, uint32_t grid, hipStream_t s) {
  const int mc = (int(a.D) + 63) / 64;
  const int npl = a.L <= 64 ? 1 : (a.L <= 128 ? 2 : 4);
#define SR(M, P) k_search<M, P><<<grid, 64, 0, s>>>(a)
#define SR_P(M)                                   \
  do {                                            \
    if (npl == 1) SR(M, 1);                       \
    else if (npl == 2) SR(M, 2);                  \
    else SR(M, 4);                                \
  } while (0)
  if (mc <= 2) SR_P(2);
  else if (mc <= 4) SR_P(4);
  else SR_P(16);
#undef SR_P
#undef SR
}

}  // namespace tg
