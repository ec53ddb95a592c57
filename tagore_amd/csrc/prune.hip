// prune.hip -- CFS (Collect-Filter-Store) pruning kernels for gfx950.
//
// One wave64 workgroup per node.  The Collect step (path beam search / two-hop list), the
// incremental Filter and the Store of reverse edges follow the deterministic rule spec that
// the CPU oracle restates (oracle/tagore_oracle.cpp), so the adjacency produced here is
// bit-identical to it:
//   collect_path  Tagore_src.cu:1209-1349     collect_2hop      Tagore_src.cu:1352-1396
//   Filter_path   Tagore_src.cu:1399-1485     rules             Tagore_src.cu:1488-1495
//   filter_reverse Tagore_src.cu:1537-1657    cal_ep_gpu        Tagore_src.cu:1889-2058
//   filter_cagra  Tagore_src.cu:1668-1710     filter_dpg        Tagore_src.cu:1713-1785
//   filter_reverse_1hop Tagore_src.cu:1812-1886
// Distances are the canonical FP32 squared L2 on FP32 rows (16 lanes per distance, four
// distances in flight per wave).  The kNN input graph is never modified: pruned rows go to
// separate buffers, so one kNN build can feed several prunings.
#include "common.hpp"
#include "kernels.hpp"

namespace tg {

constexpr int RING = 480;

// shader clock when phase profiling is on (PruneArgs::phase_clock), else 0
__device__ __forceinline__ long long pclock(const PruneArgs& a) {
  return a.phase_clock ? clock64() : 0;
}

// phase-clock sums of one workgroup (profiling): distance batches, and the beam rounds'
// candidate step, sort, merge and id-set rebuild.  Held in registers (no address taken).
struct PhaseAcc {
  long long ev = 0;
  long long sub[4] = {0, 0, 0, 0};
};
constexpr int POOL_MAX = 1024;  // K + 480 <= 608, padded to a power of two

// -------------------------------------------------------------------------------------
// Id set in LDS: open addressing over 2^bits slots, INV marks an empty slot.  Callers size
// the table above the most members it can hold, so a probe always ends.
// -------------------------------------------------------------------------------------
constexpr int HSET = 512;  // k_cal_ep: up to 256 members
__device__ __forceinline__ uint32_t hslot(uint32_t id, int bits) {
  return (id * 0x9E3779B1u) >> (32 - bits);
}
__device__ void hset_clear(uint32_t* tab, int bits) {
  for (int i = threadIdx.x; i < (1 << bits); i += TG_WAVE) tab[i] = TG_INV;
}
__device__ __forceinline__ void hset_insert(uint32_t* tab, uint32_t id, int bits) {
  if (id == TG_INV) return;
  const uint32_t mask = (1u << bits) - 1;
  uint32_t h = hslot(id, bits);
  for (;;) {
    const uint32_t old = atomicCAS(&tab[h], TG_INV, id);
    if (old == TG_INV || old == id) return;
    h = (h + 1) & mask;
  }
}
__device__ __forceinline__ bool hset_has(const uint32_t* tab, uint32_t id, int bits) {
  const uint32_t mask = (1u << bits) - 1;
  uint32_t h = hslot(id, bits);
  for (;;) {
    const uint32_t v = tab[h];
    if (v == id) return true;
    if (v == TG_INV) return false;
    h = (h + 1) & mask;
  }
}

// -------------------------------------------------------------------------------------
// Beam search shared by cal_ep and path Collect (spec in the oracle's beam_search).
// The beam lives in registers: position p = b*64 + lane is slot b of lane `lane`
// (b < W, TOPM <= 64*W; positions >= TOPM hold TG_KEY_INF).  A round evaluates the fresh
// candidates of the expanded node, sorts all K candidates in registers (S holds them by
// list position, 64*W >= K slots) and takes the TOPM smallest of beam and candidates with
// one bitonic merge: min(beam[p], cand[64W-1-p]) is bitonic and holds the 64W smallest of
// both, so sorting it gives the first 64W keys of merge(beam, candidates).
// tab: id set of the previous round's beam and candidates (the dedup set).  An entry is
// explored when its id is among the expanded ids (held in a register, one per lane): the
// beam never holds two finite entries with one id, and an expanded node that falls out of
// the beam can never re-enter it (the beam's TOPM-th key only decreases).  `out` (optional)
// receives the final beam (TOPM keys).
// -------------------------------------------------------------------------------------
template <int MC, class T, int W>
__device__ __forceinline__ void beam_search(const PruneArgs& a, const Row16<MC>& tgt,
                                            uint32_t start,
                            int max_exp, uint64_t* S, uint32_t* tab, int hbits,
                            uint16_t* cl, uint64_t* ring, uint64_t* out, uint32_t& ring_c,
                            uint32_t& n_eval, uint32_t& n_adj, PhaseAcc& ph) {
  const int lane = threadIdx.x;
  const int K = a.K, TOPM = a.TOPM, D = a.D;
  constexpr int NS = 64 * W;
  const T* X = reinterpret_cast<const T*>(a.rows);
  const uint64_t INFKEY = mk_key(__uint_as_float(0x7F800000u), 0u);
  // start list S = G[start] with distances (ring slots 0..K-1 in list order)
  for (int j = lane; j < NS; j += TG_WAVE)
    S[j] = j < K ? (INFKEY | a.G[size_t(start) * K + j]) : TG_KEY_INF;
  __syncthreads();
  eval_batch<MC, T, true>(X, D, tgt, K, [&](int q) { return key_id(S[q]); },
                 [&](int q, float d) {
                   const uint64_t k = mk_key(d, key_id(S[q]));
                   S[q] = k;
                   if (ring) ring[q % RING] = k;
                 });
  n_eval += K;
  n_adj += 1;
  ring_c = K;
  hset_clear(tab, hbits);
  __syncthreads();
  uint64_t v[W], bk[W];
#pragma unroll
  for (int r = 0; r < W; ++r) {
    v[r] = S[r * 64 + lane];
    hset_insert(tab, key_id(v[r]), hbits);  // the first dedup set is ids(S); padding ids are INV
  }
  reg_bitonic<W>(v);
#pragma unroll
  for (int b = 0; b < W; ++b) bk[b] = b * 64 + lane < TOPM ? v[b] : TG_KEY_INF;
  __syncthreads();
  // the expanded ids in registers: lane e < nexpl holds the e-th (max_exp <= 64)
  uint32_t ex = TG_INV;
  int nexpl = 0;
  for (int r = 0; r < max_exp; ++r) {
    const long long s0 = pclock(a);
    // first explorable beam entry (position order)
    int first = -1;
    uint32_t x = 0;
#pragma unroll
    for (int b = 0; b < W; ++b) {
      if (first >= 0) break;
      const uint32_t id = key_id(bk[b]);
      bool ok = key_finite(bk[b]);
      for (int e = 0; e < nexpl; ++e)
        ok = ok & (id != uint32_t(__builtin_amdgcn_readlane(int(ex), e)));
      const uint64_t m = ballot(ok);
      if (m) {
        const int fl = __ffsll((unsigned long long)m) - 1;
        first = b * 64 + fl;
        x = uint32_t(__builtin_amdgcn_readlane(int(id), fl));
      }
    }
    if (first < 0) break;
    __syncthreads();
    if (lane == nexpl) ex = x;
    ++nexpl;
    n_adj += 1;
    // new candidates: duplicates of the dedup set get +inf; the rest compacted in j order
    int cnt = 0;
#pragma unroll
    for (int base = 0; base < NS; base += TG_WAVE) {
      const int j = base + lane;
      bool fresh = false;
      if (j < K) {
        const uint32_t y = a.G[size_t(x) * K + j];
        fresh = !hset_has(tab, y, hbits);
        S[j] = INFKEY | y;
      } else {
        S[j] = TG_KEY_INF;
      }
      const uint64_t fm = ballot(fresh);
      if (fresh) cl[cnt + mask_rank(fm)] = uint16_t(j);
      cnt += __popcll(fm);
    }
    __syncthreads();
    const uint32_t c0 = ring_c;
    const long long e0 = pclock(a);
    ph.sub[0] += e0 - s0;
    eval_batch<MC, T, true>(X, D, tgt, cnt, [&](int q) { return key_id(S[cl[q]]); },
                   [&](int q, float d) {
                     const int j = cl[q];
                     const uint64_t k = mk_key(d, key_id(S[j]));
                     S[j] = k;
                     if (ring) ring[(c0 + q) % RING] = k;
                   });
    ph.ev += pclock(a) - e0;
    ring_c += cnt;
    n_eval += cnt;
    __syncthreads();
    const long long s1 = pclock(a);
#pragma unroll
    for (int q = 0; q < W; ++q) v[q] = S[q * 64 + lane];
    // next dedup set: ids of this round's beam and candidates
    hset_clear(tab, hbits);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < W; ++q) {
      hset_insert(tab, key_id(bk[q]), hbits);
      hset_insert(tab, key_id(v[q]), hbits);
    }
    const long long s2 = pclock(a);
    // Only candidates below the beam's TOPM-th key can enter it (a fresh id is in neither
    // the beam nor the previous candidates, so no key ties one in the beam).  With two
    // slots per lane and at most 64 such candidates (and TOPM <= 64), they are compacted
    // and merged with one-slot sorts; none leaves the beam as it is.  Either way the beam
    // becomes the TOPM smallest keys of beam and candidates.
    int m = 64 * W;
    bool small = false;
    if (W > 1 && TOPM <= 64) {
      const uint64_t wk = bk[0];
      const uint64_t worst = (uint64_t(uint32_t(__builtin_amdgcn_readlane(int(wk >> 32), TOPM - 1))) << 32) |
                             uint32_t(__builtin_amdgcn_readlane(int(uint32_t(wk)), TOPM - 1));
      uint64_t qm[W];
      m = 0;
#pragma unroll
      for (int q = 0; q < W; ++q) {
        qm[q] = ballot(key_finite(v[q]) && v[q] < worst);
        m += __popcll(qm[q]);
      }
      small = m <= 64;
      if (small && m > 0) {
        int base = 0;
#pragma unroll
        for (int q = 0; q < W; ++q) {
          if ((qm[q] >> lane) & 1ull) S[base + mask_rank(qm[q])] = v[q];
          base += __popcll(qm[q]);
        }
        if (lane >= m) S[lane] = TG_KEY_INF;
        __syncthreads();
        v[0] = S[lane];
      }
    }
    long long s3 = 0;
    if (small) {
      if (m > 0) {
        // keys in lanes 0..m-1, +inf above: sorting blocks of the next power of two is enough
        uint64_t w1 = v[0];
        if (m <= 16) reg_bitonic_upto<16>(w1);
        else if (m <= 32) reg_bitonic_upto<32>(w1);
        else reg_bitonic_upto<64>(w1);
        s3 = pclock(a);
        const uint64_t rv = lane_xor64(w1, 63);
        uint64_t c[1] = {bk[0] < rv ? bk[0] : rv};
        reg_bitonic_merge<1>(c);
        bk[0] = lane < TOPM ? c[0] : TG_KEY_INF;
      } else {
        s3 = pclock(a);
      }
    } else {
      reg_bitonic<W>(v);
      s3 = pclock(a);
      uint64_t c[W];
#pragma unroll
      for (int b = 0; b < W; ++b) {
        const uint64_t rv = lane_xor64(v[W - 1 - b], 63);  // candidate 64W-1-p
        c[b] = bk[b] < rv ? bk[b] : rv;
      }
      reg_bitonic_merge<W>(c);
#pragma unroll
      for (int b = 0; b < W; ++b) bk[b] = b * 64 + lane < TOPM ? c[b] : TG_KEY_INF;
    }
    __syncthreads();
    const long long s4 = pclock(a);
    ph.sub[1] += s3 - s2;
    ph.sub[2] += s4 - s3;
    ph.sub[3] += s2 - s1;
  }
  if (out) {
#pragma unroll
    for (int b = 0; b < W; ++b)
      if (b * 64 + lane < TOPM) out[b * 64 + lane] = bk[b];
  }
  __syncthreads();
}

// -------------------------------------------------------------------------------------
// The reverse pass's pool: the forward row (deg keys, ascending: selections come in key
// order) and the reverse top (nr keys, ascending) as one sorted sequence of 64 NPL keys,
// +inf after L = deg + nr.  [forward | +inf | reverse, last first] is bitonic, so one
// bitonic merge gives what a sort of the union gives.
// -------------------------------------------------------------------------------------
template <int NPL>
__device__ __forceinline__ void merge_pool(const PruneArgs& a, uint32_t u, int deg,
                                           const uint64_t* rev, int nr, uint64_t* P) {
  const int lane = threadIdx.x;
  constexpr int n = 64 * NPL;
  uint64_t v[NPL];
#pragma unroll
  for (int b = 0; b < NPL; ++b) {
    const int i = b * 64 + lane;
    v[b] = TG_KEY_INF;
    if (i < deg) v[b] = mk_key(a.fwd_d[size_t(u) * a.FD + i], a.fwd_ids[size_t(u) * a.FD + i]);
    else if (i >= n - nr) v[b] = rev[n - 1 - i];
  }
  reg_bitonic_merge<NPL>(v);
#pragma unroll
  for (int b = 0; b < NPL; ++b) P[b * 64 + lane] = v[b];
  __syncthreads();
}

// -------------------------------------------------------------------------------------
// In a sorted pool of n keys, replace the id of every entry whose id equals its
// predecessor's (original values) by INV.
// -------------------------------------------------------------------------------------
__device__ void mark_dups(uint64_t* P, int n) {
  const int lane = threadIdx.x;
  bool dup[POOL_MAX / TG_WAVE];
#pragma unroll
  for (int q = 0; q < POOL_MAX / TG_WAVE; ++q) {
    const int i = q * TG_WAVE + lane;
    dup[q] = (i < n && i > 0) ? (key_id(P[i]) == key_id(P[i - 1])) : false;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < POOL_MAX / TG_WAVE; ++q) {
    const int i = q * TG_WAVE + lane;
    if (dup[q]) P[i] = key_with_id(P[i], TG_INV);
  }
  __syncthreads();
}

// first index >= from, < L, whose id is not INV (and not `skip`); L if none
__device__ int next_valid(const uint64_t* P, int from, int L, uint32_t skip) {
  const int lane = threadIdx.x;
  for (int base = from; base < L; base += TG_WAVE) {
    const int i = base + lane;
    bool ok = false;
    if (i < L) {
      const uint32_t id = key_id(P[i]);
      ok = id != TG_INV && id != skip;
    }
    const uint64_t m = ballot(ok);
    if (m) return base + __ffsll((unsigned long long)m) - 1;
  }
  return L;
}

// -------------------------------------------------------------------------------------
// The next selection and the next round's candidate list, in one pass over the entries
// cl[from..from+n) (pool indices, in pool order): the first entry not invalidated is the
// selection (its pool index goes to pi, -1 when there is none); the entries after it that
// are not invalidated are the candidates, compacted to cl[0..), except those repeating the
// selection's id, which are invalidated uncounted (a round's start rule).  Returns the
// candidates' number.  Each 64-entry chunk is read before any of its entries is written,
// and writes land below the next chunk, so the compaction is in place.
// -------------------------------------------------------------------------------------
__device__ __forceinline__ int advance_cl(uint64_t* P, uint16_t* cl, int from, int n, int& pi) {
  const int lane = threadIdx.x;
  int out = 0;
  uint32_t sid = TG_INV;
  pi = -1;
  for (int base = 0; base < n; base += TG_WAVE) {
    const int j = base + lane;
    const int c = j < n ? int(cl[from + j]) : 0;
    const uint32_t id = j < n ? key_id(P[c]) : TG_INV;
    bool k = id != TG_INV;
    if (pi < 0) {
      const uint64_t m = ballot(k);
      if (m) {
        const int f = __ffsll((unsigned long long)m) - 1;
        pi = __builtin_amdgcn_readlane(c, f);
        sid = uint32_t(__builtin_amdgcn_readlane(int(id), f));
        k = k && lane > f;
      }
    }
    if (k && id == sid) {
      P[c] = key_with_id(P[c], TG_INV);
      k = false;
    }
    const uint64_t m = ballot(k);
    __syncthreads();
    if (k) cl[out + mask_rank(m)] = uint16_t(c);
    out += __popcll(m);
    __syncthreads();
  }
  return out;
}

// -------------------------------------------------------------------------------------
// Incremental Filter over the sorted pool P[0..L) (spec A.4 in the oracle).  Writes the
// kept entries to sel[] and returns their number.  cl holds the round's candidates: the pool
// indices after the current selection p that are not invalidated and do not repeat p's id,
// in pool order, so a round touches only the survivors of the previous round.
// -------------------------------------------------------------------------------------
template <int MC, class T>
__device__ __forceinline__ int filter_pool(const PruneArgs& a, uint64_t* P, int L, uint32_t u,
                                           int rule, uint64_t* sel, uint16_t* cl,
                                           uint32_t& n_eval, PhaseAcc& ph) {
  const int lane = threadIdx.x, t = lane & 15;
  const int D = a.D, FD = a.FD;
  int tix = next_valid(P, 0, L, u);
  if (tix >= L) return 0;
  int nsel = 1;
  if (lane == 0) sel[0] = P[tix];
  float cur = key_d(P[tix]);
  uint32_t pid = key_id(P[tix]);
  ++tix;
  // the first round's candidates; entries repeating p's id are invalidated uncounted
  int nl = 0;
  for (int base = tix; base < L; base += TG_WAVE) {
    const int i = base + lane;
    const uint32_t id = i < L ? key_id(P[i]) : TG_INV;
    bool live = id != TG_INV;
    if (live && id == pid) {
      P[i] = key_with_id(P[i], TG_INV);
      live = false;
    }
    const uint64_t m = ballot(live);
    if (live) cl[nl + mask_rank(m)] = uint16_t(i);
    nl += __popcll(m);
  }
  __syncthreads();
  while (nsel < FD) {
    const int cnt = nl;  // this round's candidates cl[0..cnt)
    Row16<MC> prow;
    const long long e0 = pclock(a);
    if constexpr (std::is_same<T, uint8_t>::value) {
      // Two rounds in one pass: the first candidate c1 is the next selection exactly when
      // this round keeps it, which d(p, c1) -- two rows in registers -- decides.  Then each
      // candidate row is read once for d(p, c) and d(c1, c), and every candidate meets the
      // two rounds' rules in their order: one invalidated by p is not an evaluation of the
      // second round, an entry repeating c1's id is dropped uncounted (as at a round's
      // start), the others count for the second round.
      const uint8_t* X8 = reinterpret_cast<const uint8_t*>(a.rows);
      Row16<MC> qrow;
      const bool two = cnt > 1 && nsel + 1 < FD;
      const int i1 = two ? int(cl[0]) : 0;
      row16_u8_load<MC>(prow, X8 + size_t(pid) * D, D);
      if (two) row16_u8_load<MC>(qrow, X8 + size_t(key_id(P[i1])) * D, D);
      row16_u8_norm<MC>(prow);
      if (two) {
        row16_u8_norm<MC>(qrow);
        const uint64_t k1 = P[i1];
        const uint32_t id1 = key_id(k1);
        const float cur1 = key_d(k1);
        const float d1 = dist_rows_u8<MC>(prow, qrow, D);
        if (!(rule_eval(rule, d1, cur1, cur, a.th) < 0.0f) && i1 + 1 < L) {
          int n2 = 0;
          eval_batch2_u8<MC>(X8, D, prow, qrow, cnt, [&](int q) { return key_id(P[cl[q]]); },
                             [&](int q, float da, float db) {
                               if (q == 0) return;  // c1: kept by this round
                               const int i = cl[q];
                               const uint64_t k = P[i];
                               if (rule_eval(rule, da, key_d(k), cur, a.th) < 0.0f) {
                                 P[i] = key_with_id(k, TG_INV);
                               } else if (key_id(k) == id1) {
                                 P[i] = key_with_id(k, TG_INV);
                               } else {
                                 ++n2;
                                 if (rule_eval(rule, db, key_d(k), cur1, a.th) < 0.0f)
                                   P[i] = key_with_id(k, TG_INV);
                               }
                             });
#pragma unroll
          for (int o = 32; o > 0; o >>= 1) n2 += __shfl_xor(n2, o, TG_WAVE);
          ph.ev += pclock(a) - e0;
          n_eval += cnt + n2;
          __syncthreads();
          // the first round's selection is c1; the second's is its first survivor
          if (lane == 0) sel[nsel] = k1;
          ++nsel;
          int pi;
          nl = advance_cl(P, cl, 1, cnt - 1, pi);
          if (pi >= 0) {
            if (lane == 0) sel[nsel] = P[pi];
            cur = key_d(P[pi]);
            pid = key_id(P[pi]);
            tix = pi;
          } else {
            tix = L;
          }
          ++nsel;
          ++tix;
          __syncthreads();
          if (tix >= L) break;
          continue;
        }
      }
    } else {
      load_row16<MC>(prow, reinterpret_cast<const T*>(a.rows) + size_t(pid) * D, D, t);
    }
    eval_batch<MC>(reinterpret_cast<const T*>(a.rows), D, prow, cnt, [&](int q) { return key_id(P[cl[q]]); },
                   [&](int q, float res) {
                     const int i = cl[q];
                     const uint64_t k = P[i];
                     if (rule_eval(rule, res, key_d(k), cur, a.th) < 0.0f)
                       P[i] = key_with_id(k, TG_INV);
                   });
    ph.ev += pclock(a) - e0;
    n_eval += cnt;
    __syncthreads();
    // the next selection is the first survivor; the survivors after it are the next round's
    // candidates
    int pi;
    nl = advance_cl(P, cl, 0, cnt, pi);
    if (pi >= 0) {
      if (lane == 0) sel[nsel] = P[pi];
      cur = key_d(P[pi]);
      pid = key_id(P[pi]);
      tix = pi;
    } else {
      tix = L;
    }
    ++nsel;  // an exhausted pool appends a placeholder, dropped below
    ++tix;
    __syncthreads();
    if (tix >= L) break;
  }
  return nsel - 1;  // the last appended entry is never kept (Tagore_src.cu:1462-1477)
}

__device__ __forceinline__ void count_add(unsigned long long* c, int idx, uint32_t v) {
  if (c && threadIdx.x == 0 && v) atomicAdd(&cnt_slot(c)[idx], (unsigned long long)v);
}

// -------------------------------------------------------------------------------------
// Forward pass (Collect + Filter + Store reverse pushes).  kind 0: path, 1: two-hop.
//
// LDS (dynamic, sized from K, TOPM and FD so that more workgroups fit a CU):
//   during the beam: the 480-slot ring (padded to 512 for its sort), then S | tab | beam
//       compaction indices;
//   afterwards: the pool (K + 480 keys, over the ring and the dead start of the beam's
//       scratch), then the selected list and the filter's compaction indices (K + 480).
// The pool is the merge of the sorted own list and the sorted ring, formed in registers as
// one bitonic merge of [own list ascending | +inf | ring descending] (1024 slots); it is
// the sequence a sort of their union gives.
// -------------------------------------------------------------------------------------
constexpr int RING_P = 512;

struct FwdLayout {
  int W, NS, L, HB, offB, offF, bytes;
};
__host__ __device__ inline FwdLayout fwd_layout(int K, int TOPM, int FD) {
  FwdLayout l;
  l.W = (K <= 64 && TOPM <= 64) ? 1 : 2;
  l.NS = 64 * l.W;
  l.L = K + RING;
  // the beam's id set holds at most TOPM + K ids: 256 slots up to 192 of them (three
  // quarters full at most), else 512
  l.HB = TOPM + K <= 192 ? 8 : 9;
  l.offB = RING_P * 8;
  l.offF = ((l.L > RING_P ? l.L : RING_P) * 8 + 15) / 16 * 16;
  const int beam = l.offB + l.NS * 8 + (4 << l.HB) + l.NS * 2;
  const int filt = l.offF + (FD + 1) * 8 + l.L * 2;
  l.bytes = ((beam > filt ? beam : filt) + 15) / 16 * 16;
  return l;
}

template <int MC, class T, int KIND, int W>
// Six waves per SIMD: 80 VGPRs at most (the 64-candidate distance steps need 86 otherwise).
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6))) void k_forward(PruneArgs a) {
  extern __shared__ __attribute__((aligned(16))) unsigned char dsm[];
  const uint32_t u = row_of_block(a.perm, a.N);
  if (u >= a.N) return;
  const int lane = threadIdx.x, t = lane & 15;
  const int K = a.K, D = a.D;
  const FwdLayout ly = fwd_layout(K, int(a.TOPM), int(a.FD));
  uint64_t* R = reinterpret_cast<uint64_t*>(dsm);
  uint64_t* P = R;  // after the pool is formed
  uint64_t* S = reinterpret_cast<uint64_t*>(dsm + ly.offB);
  uint32_t* tab = reinterpret_cast<uint32_t*>(S + ly.NS);
  uint16_t* clb = reinterpret_cast<uint16_t*>(tab + (1 << ly.HB));
  uint64_t* sel = reinterpret_cast<uint64_t*>(dsm + ly.offF);  // after the beam
  uint16_t* cl = reinterpret_cast<uint16_t*>(sel + a.FD + 1);
  const int L = ly.L;
  Row16<MC> xu;
  load_row16<MC>(xu, reinterpret_cast<const T*>(a.rows) + size_t(u) * D, D, t);
  for (int i = lane; i < RING_P; i += TG_WAVE)
    R[i] = i < RING ? mk_key(__uint_as_float(0x7F800000u), TG_INV) : TG_KEY_INF;
  __syncthreads();
  uint32_t n_col = 0, n_adj = 0, n_flt = 0;
  PhaseAcc ph;
  const long long c0 = pclock(a);
  if (KIND == 0) {
    uint32_t ring_c = 0;
    const uint32_t ep = a.ep[a.seg_of ? a.seg_of[u] : 0];
    beam_search<MC, T, W>(a, xu, ep, 10, S, tab, ly.HB, clb, R, nullptr, ring_c, n_col,
                          n_adj, ph);
  } else {
    // two-hop ids y_i = G[G[u][i/K]][i%K], i < 480
    for (int i = lane; i < RING; i += TG_WAVE) {
      const uint32_t v = a.G[size_t(u) * K + i / K];
      R[i] = mk_key(0.f, a.G[size_t(v) * K + (i % K)]);
    }
    __syncthreads();
    eval_batch<MC>(reinterpret_cast<const T*>(a.rows), D, xu, RING, [&](int q) { return key_id(R[q]); },
                   [&](int q, float d) { R[q] = mk_key(d, key_id(R[q])); });
    n_col = RING;
    n_adj = (RING + K - 1) / K;
  }
  __syncthreads();
  const long long c1 = pclock(a);
  // pool = merge(own list, sorted ring), in registers; the own list is a kNN row, sorted by
  // (distance, id) already (builds keep them so, tg_set_knn rejects others)
  uint64_t pv[16];
  {
    uint64_t o[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int j = r * 64 + lane;
      o[r] = j < K ? mk_key(a.GD[size_t(u) * K + j], a.G[size_t(u) * K + j]) : TG_KEY_INF;
    }
    uint64_t rr[RING_P / 64];
#pragma unroll
    for (int q = 0; q < RING_P / 64; ++q) rr[q] = R[q * 64 + lane];
    reg_bitonic<RING_P / 64, uint64_t, true>(rr);
    pv[0] = o[0];
    pv[1] = o[1];
#pragma unroll
    for (int q = 2; q < 8; ++q) pv[q] = TG_KEY_INF;
#pragma unroll
    for (int q = 0; q < 8; ++q) pv[8 + q] = rr[q];
  }
  const long long c1s = pclock(a);
  reg_bitonic_merge<16>(pv);
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 16; ++q)
    if (q * 64 + lane < L) P[q * 64 + lane] = pv[q];
  __syncthreads();
  // entries repeating their predecessor's id get id INV
  constexpr int PQ = (128 + RING + TG_WAVE - 1) / TG_WAVE;
  bool dup[PQ];
#pragma unroll
  for (int q = 0; q < PQ; ++q) {
    const int i = q * TG_WAVE + lane;
    dup[q] = (i < L && i > 0) ? (key_id(P[i]) == key_id(P[i - 1])) : false;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PQ; ++q)
    if (dup[q]) P[q * TG_WAVE + lane] = key_with_id(P[q * TG_WAVE + lane], TG_INV);
  __syncthreads();
  const long long c2 = pclock(a);
  const int deg = filter_pool<MC, T>(a, P, L, u, a.fwd_rule, sel, cl, n_flt, ph);
  __syncthreads();
  if (a.phase_clock && lane == 0 && a.counters) {
    const long long c3 = clock64();
    atomicAdd(&cnt_slot(a.counters)[6], (unsigned long long)(c1 - c0));
    atomicAdd(&cnt_slot(a.counters)[7], (unsigned long long)(c3 - c2));
    atomicAdd(&cnt_slot(a.counters)[8], (unsigned long long)(c2 - c1));
    atomicAdd(&cnt_slot(a.counters)[9], (unsigned long long)ph.ev);
    for (int i = 0; i < 4; ++i) atomicAdd(&cnt_slot(a.counters)[10 + i], (unsigned long long)ph.sub[i]);
    atomicAdd(&cnt_slot(a.counters)[14], (unsigned long long)(c1s - c1));
  }
  for (int k = lane; k < deg; k += TG_WAVE) {
    const uint64_t s = sel[k];
    a.fwd_ids[size_t(u) * a.FD + k] = key_id(s);
    a.fwd_d[size_t(u) * a.FD + k] = key_d(s);
    pb_push(a.rev, key_id(s), mk_key(key_d(s), u));
  }
  if (lane == 0) a.fwd_deg[u] = deg;
  count_add(a.counters, 0, n_col);
  count_add(a.counters, 1, n_flt);
  count_add(a.counters, 3, n_adj);
}

// -------------------------------------------------------------------------------------
// Reverse pass (filter_reverse, spec A.7): forward list plus the K best reverse pushes;
// re-filtered with the reverse rule when more than FD, otherwise their union.
// -------------------------------------------------------------------------------------
template <int MC, class T>
__global__ __launch_bounds__(64) void k_reverse(PruneArgs a) {
  __shared__ uint64_t buf[256];
  __shared__ uint64_t P[256];
  __shared__ uint16_t cl[256];
  __shared__ uint64_t sel[129];
  const uint32_t u = row_of_block(a.perm, a.N);
  if (u >= a.N) return;
  const int lane = threadIdx.x;
  const int K = a.K, FD = a.FD;
  const int nr = pb_top(a.rev, u, K, false, [](uint64_t) { return true; }, buf);
  const int deg = a.fwd_deg[u];
  const int L = deg + nr;
  uint32_t* row = a.out_ids + size_t(u) * FD;
  uint32_t n_flt = 0;
  if (L > FD) {
    const int n = pow2_at_least(max(L, TG_WAVE));  // L <= FD + K <= 256
    if (n == 64) merge_pool<1>(a, u, deg, buf, nr, P);
    else if (n == 128) merge_pool<2>(a, u, deg, buf, nr, P);
    else merge_pool<4>(a, u, deg, buf, nr, P);
    mark_dups(P, n);
    PhaseAcc ph;  // not reported for the reverse pass
    const int nd = filter_pool<MC, T>(a, P, L, u, a.rev_rule, sel, cl, n_flt, ph);
    __syncthreads();
    for (int k = lane; k < FD; k += TG_WAVE) row[k] = k < nd ? key_id(sel[k]) : TG_INV;
    if (lane == 0) a.out_deg[u] = nd;
  } else {
    // union: forward ids in order, then reverse ids (R order) not already present
    const uint32_t* f = a.fwd_ids + size_t(u) * FD;
    for (int k = lane; k < deg; k += TG_WAVE) row[k] = f[k];
    bool add = false;
    uint32_t w = 0;
    if (lane < nr) {
      w = key_id(buf[lane]);
      add = true;
      for (int j = 0; j < deg; ++j) add = add && (f[j] != w);
    }
    // nr <= L <= FD <= 128; handle lanes 64..127 in a second step
    const uint64_t m0 = ballot(add);
    if (add) row[deg + mask_rank(m0)] = w;
    int n = deg + __popcll(m0);
    if (nr > TG_WAVE) {
      bool add2 = false;
      uint32_t w2 = 0;
      if (TG_WAVE + lane < nr) {
        w2 = key_id(buf[TG_WAVE + lane]);
        add2 = true;
        for (int j = 0; j < deg; ++j) add2 = add2 && (f[j] != w2);
      }
      const uint64_t m1 = ballot(add2);
      if (add2) row[n + mask_rank(m1)] = w2;
      n += __popcll(m1);
    }
    for (int k = n + lane; k < FD; k += TG_WAVE) row[k] = TG_INV;
    if (lane == 0) a.out_deg[u] = n;
  }
  count_add(a.counters, 2, n_flt);
}

// -------------------------------------------------------------------------------------
// One-hop rank (CAGRA: id membership; DPG: occlusion by distance), spec A.8 / A.9.
// -------------------------------------------------------------------------------------
template <int MC, class T, int DPG>
__global__ __launch_bounds__(64) void k_rank_forward(PruneArgs a) {
  __shared__ uint64_t srt[256];   // (id << 32 | pos) sorted, then (detour << 32 | id)
  __shared__ uint32_t det[128];
  __shared__ uint32_t nei[128];
  const uint32_t u = row_of_block(a.perm, a.N);
  if (u >= a.N) return;
  const int lane = threadIdx.x, g = lane >> 4, t = lane & 15;
  const int K = a.K, D = a.D, d = a.FD;
  const int n = pow2_at_least(K);
  for (int j = lane; j < 128; j += TG_WAVE) {
    det[j] = 0;
    nei[j] = j < K ? a.G[size_t(u) * K + j] : TG_INV;
  }
  __syncthreads();
  if (!DPG) {
    for (int j = lane; j < n; j += TG_WAVE)
      srt[j] = j < K ? ((uint64_t(nei[j]) << 32) | uint32_t(j)) : TG_KEY_INF;
    __syncthreads();
    wave_sort<128>(srt, n);
    for (int r = 0; r < K; ++r) {
      const uint32_t v = nei[r];
      for (int l = lane; l < K; l += TG_WAVE) {
        const uint32_t y = a.G[size_t(v) * K + l];
        int lo = 0, hi = K;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (uint32_t(srt[mid] >> 32) < y) lo = mid + 1; else hi = mid;
        }
        if (lo < K && uint32_t(srt[lo] >> 32) == y) {
          const int j = int(uint32_t(srt[lo]));
          if (j > r) atomicAdd(&det[j], 1u);
        }
      }
    }
    count_add(a.counters, 3, K);
  } else {
    // pairs r < j, row by row; the four groups take every fourth j of a row
    const int npairs = K * (K - 1) / 2;
    for (int r = 0; r + 1 < K; ++r) {
      const T* xr = reinterpret_cast<const T*>(a.rows) + size_t(nei[r]) * D;
      for (int j = r + 1 + g; j < K; j += 4) {
        const float dd =
            dist16_mm<MC>(xr, reinterpret_cast<const T*>(a.rows) + size_t(nei[j]) * D, D, t);
        if (t == 0 && dd < a.GD[size_t(u) * K + j]) atomicAdd(&det[j], 1u);
      }
    }
    count_add(a.counters, 1, npairs);
  }
  __syncthreads();
  for (int j = lane; j < n; j += TG_WAVE)
    srt[j] = j < K ? ((uint64_t(det[j]) << 32) | nei[j]) : TG_KEY_INF;
  __syncthreads();
  wave_sort<128>(srt, n);
  for (int i = lane; i < d; i += TG_WAVE) {
    const uint64_t k = srt[i];
    a.fwd_ids[size_t(u) * d + i] = uint32_t(k);
    pb_push(a.rev, uint32_t(k), (k & 0xFFFFFFFF00000000ull) | u);
  }
  if (lane == 0) a.fwd_deg[u] = d;
}

// filter_reverse_1hop (spec A.10): half forward, up to half reverse, forward fill.
__global__ __launch_bounds__(64) void k_rank_reverse(PruneArgs a) {
  __shared__ uint64_t buf[256];
  __shared__ uint32_t A[128];
  const uint32_t u = blockIdx.x;
  if (u >= a.N) return;
  const int lane = threadIdx.x;
  const int d = a.FD, h = d / 2;
  // R1[u]: the d smallest (detour, src) pushes, already sorted by (detour, id)
  const int r = pb_top(a.rev, u, d, false, [](uint64_t) { return true; }, buf);
  const uint32_t* T = a.fwd_ids + size_t(u) * d;
  uint32_t* row = a.out_ids + size_t(u) * d;
  // Rv sorted by (id, detour); later copies of an id get detour INV; then by (detour,id).
  // Pushes carry distinct sources, so the relabelling never fires; it is kept for parity.
  if (lane == 0) {
    // insertion-sort by id on buf[128..128+r)
    for (int i = 0; i < r; ++i) {
      const uint64_t k = (uint64_t(uint32_t(buf[i])) << 32) | uint32_t(buf[i] >> 32);
      int j = i;
      while (j > 0 && buf[128 + j - 1] > k) {
        buf[128 + j] = buf[128 + j - 1];
        --j;
      }
      buf[128 + j] = k;
    }
    for (int i = r - 1; i >= 1; --i)
      if (uint32_t(buf[128 + i] >> 32) == uint32_t(buf[128 + i - 1] >> 32))
        buf[128 + i] = (buf[128 + i] & 0xFFFFFFFF00000000ull) | TG_INV;
    // back to (detour << 32 | id) order
    for (int i = 0; i < r; ++i) {
      const uint64_t k = (uint64_t(uint32_t(buf[128 + i])) << 32) | uint32_t(buf[128 + i] >> 32);
      int j = i;
      while (j > 0 && buf[j - 1] > k) {
        buf[j] = buf[j - 1];
        --j;
      }
      buf[j] = k;
    }
    int na = 0;
    const int lim = (r < h) ? (d - r) : h;
    for (int i = 0; i < r; ++i) {
      if (r >= h && na >= h) break;
      const uint32_t w = uint32_t(buf[i]);
      bool in = false;
      for (int j = 0; j < lim; ++j) in = in || (T[j] == w);
      if (!in) A[na++] = w;
    }
    for (int i = 0; i < h; ++i) row[i] = T[i];
    for (int i = 0; i < na; ++i) row[h + i] = A[i];
    for (int i = 0; h + na + i < d; ++i) row[h + na + i] = T[h + i];
    a.out_deg[u] = d;
  }
}

// -------------------------------------------------------------------------------------
// Entry point (cal_ep_gpu, Tagore_src.cu:1889-2058): beam of TOPM from the first node's list
// toward the centroid; one workgroup per segment.
// -------------------------------------------------------------------------------------
template <int MC, class T>
__global__ __launch_bounds__(64) void k_cal_ep(PruneArgs a, const float* centroid, uint32_t* ep) {
  __shared__ uint64_t M0[128], S[128];
  __shared__ uint32_t tab[HSET];
  __shared__ uint16_t cl[128];
  const uint32_t seg = blockIdx.x;
  const uint32_t start = a.seg_lo ? a.seg_lo[seg] : 0u;
  Row16<MC> c;
  load_row16<MC>(c, centroid + size_t(seg) * a.D, a.D, threadIdx.x & 15);
  uint32_t rc = 0, ne = 0, na = 0;
  PhaseAcc ph;  // not reported for the entry point
  if (a.K <= 64 && a.TOPM <= 64)
    beam_search<MC, T, 1>(a, c, start, 30, S, tab, 9, cl, nullptr, M0, rc, ne, na, ph);
  else
    beam_search<MC, T, 2>(a, c, start, 30, S, tab, 9, cl, nullptr, M0, rc, ne, na, ph);
  if (threadIdx.x == 0) ep[seg] = key_id(M0[0]);
  count_add(a.counters, 4, ne);
}

// Deterministic centroid: FP64 sums over chunks of 1024 rows, then over chunks in order.
// The centroid's FP64 sums in index order (chunks of 1024 rows, then the chunk partials):
// each step loads 16 values together and adds them in order, so 16 loads are in flight per
// lane and the sums are the serial ones.
__global__ void k_centroid_part(const float* X, uint32_t N, uint32_t D, double* part) {
  const uint32_t j = blockIdx.y * blockDim.x + threadIdx.x;
  if (j >= D) return;
  const size_t lo = size_t(blockIdx.x) * 1024, hi = min(size_t(N), lo + 1024);
  double s = 0.0;
  size_t i = lo;
  for (; i + 16 <= hi; i += 16) {
    float v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = X[(i + q) * D + j];
#pragma unroll
    for (int q = 0; q < 16; ++q) s += double(v[q]);
  }
  for (; i < hi; ++i) s += double(X[i * D + j]);
  part[size_t(blockIdx.x) * D + j] = s;
}
__global__ void k_centroid_final(const double* part, uint32_t nch, uint32_t N, uint32_t D,
                                 float* c) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= D) return;
  double s = 0.0;
  uint32_t ch = 0;
  for (; ch + 16 <= nch; ch += 16) {
    double v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = part[size_t(ch + q) * D + j];
#pragma unroll
    for (int q = 0; q < 16; ++q) s += v[q];
  }
  for (; ch < nch; ++ch) s += part[size_t(ch) * D + j];
  c[j] = float(s / double(N));
}

// ------------------------------------------------------------------------- launchers
template <int MC, class T>
static void launch_prune_mc(const PruneArgs& a, int phase, const float* centroid, uint32_t* ep,
                            hipStream_t s, uint32_t nseg) {
  const dim3 grid(grid_rows(a.perm, a.N)), block(64);
  switch (phase) {
    case PH_EP: k_cal_ep<MC, T><<<nseg, 64, 0, s>>>(a, centroid, ep); break;
    case PH_FWD_PATH:
      if (fwd_layout(a.K, a.TOPM, a.FD).W == 1)
        k_forward<MC, T, 0, 1><<<grid, block, fwd_layout(a.K, a.TOPM, a.FD).bytes, s>>>(a);
      else
        k_forward<MC, T, 0, 2><<<grid, block, fwd_layout(a.K, a.TOPM, a.FD).bytes, s>>>(a);
      break;
    case PH_FWD_2HOP:
      k_forward<MC, T, 1, 1><<<grid, block, fwd_layout(a.K, a.TOPM, a.FD).bytes, s>>>(a);
      break;
    case PH_REV: k_reverse<MC, T><<<grid, block, 0, s>>>(a); break;
    case PH_CAGRA: k_rank_forward<MC, T, 0><<<grid, block, 0, s>>>(a); break;
    case PH_DPG: k_rank_forward<MC, T, 1><<<grid, block, 0, s>>>(a); break;
    case PH_RANK_REV: k_rank_reverse<<<grid, block, 0, s>>>(a); break;
    default: break;
  }
}

template <class T>
static void launch_prune_t(const PruneArgs& a, int phase, const float* centroid, uint32_t* ep,
                           hipStream_t s, uint32_t nseg) {
  const int mc = (a.D + 63) / 64;
  if (mc <= 2) launch_prune_mc<2, T>(a, phase, centroid, ep, s, nseg);
  else if (mc <= 4) launch_prune_mc<4, T>(a, phase, centroid, ep, s, nseg);
  else launch_prune_mc<16, T>(a, phase, centroid, ep, s, nseg);
}

// The passes whose targets are rows of X, on the u8 copy.
template <int MC>
static void launch_prune_u8_mc(const PruneArgs& a, int phase, hipStream_t s) {
  const dim3 grid(grid_rows(a.perm, a.N)), block(64);
  const FwdLayout f = fwd_layout(a.K, a.TOPM, a.FD);
  switch (phase) {
    case PH_FWD_PATH:
      if (f.W == 1) k_forward<MC, uint8_t, 0, 1><<<grid, block, f.bytes, s>>>(a);
      else k_forward<MC, uint8_t, 0, 2><<<grid, block, f.bytes, s>>>(a);
      break;
    case PH_FWD_2HOP: k_forward<MC, uint8_t, 1, 1><<<grid, block, f.bytes, s>>>(a); break;
    case PH_REV: k_reverse<MC, uint8_t><<<grid, block, 0, s>>>(a); break;
    default: break;
  }
}

void launch_prune(const PruneArgs& a, int phase, const float* centroid, uint32_t* ep,
                  hipStream_t s, uint32_t nseg) {
  // (rows up to 256 values: wider u8 batches would not stay in registers)
  if (a.rows8 && a.D <= 256 && (phase == PH_FWD_PATH || phase == PH_FWD_2HOP || phase == PH_REV)) {
    PruneArgs b = a;
    b.rows = a.rows8;
    b.rows_bf16 = 0;
    if (a.D <= 128) launch_prune_u8_mc<2>(b, phase, s);
    else launch_prune_u8_mc<4>(b, phase, s);
    return;
  }
  if (a.rows_bf16) launch_prune_t<uint16_t>(a, phase, centroid, ep, s, nseg);
  else launch_prune_t<float>(a, phase, centroid, ep, s, nseg);
}

void launch_centroid(const float* X, uint32_t N, uint32_t D, double* part, float* c,
                     hipStream_t s) {
  const uint32_t nch = (N + 1023) / 1024;
  k_centroid_part<<<dim3(nch, (D + 63) / 64), 64, 0, s>>>(X, N, D, part);
  k_centroid_final<<<(D + 63) / 64, 64, 0, s>>>(part, nch, N, D, c);
}

}  // namespace tg
