// kernels.hpp -- argument blocks and launchers shared between the HIP kernels and the
// host orchestration (api.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"

namespace tg {

enum PrunePhase {
  PH_EP = 0,
  PH_FWD_PATH = 1,
  PH_FWD_2HOP = 2,
  PH_REV = 3,
  PH_CAGRA = 4,
  PH_DPG = 5,
  PH_RANK_REV = 6,
};

struct PruneArgs {
  const float* X;
  uint32_t N, D, K;
  const uint32_t* G;   // kNN ids [N*K], rows sorted by (d, id)
  const float* GD;     // kNN distances [N*K]
  uint32_t FD, TOPM;
  float th;
  int fwd_rule, rev_rule;  // 0 dist, 1 angle
  const uint32_t* ep;      // entry point of each segment (one entry without segments)
  uint32_t* fwd_deg;       // [N]
  uint32_t* fwd_ids;       // [N*FD]
  float* fwd_d;            // [N*FD]
  uint32_t* out_deg;       // [N]
  uint32_t* out_ids;       // [N*FD]
  PushBuf rev;             // reverse pushes
  unsigned long long* counters;  // [8] evaluation counters or null
  // rows the distances are read from: X itself, or its BF16 copy when every value of the
  // data set is exactly representable in BF16 (identical FP32 values, half the bytes)
  const void* rows;
  int rows_bf16;
  // u8 copy of X (or null) when every value is an integer in [0, 255] and every squared
  // distance is below 2^24: the forward and reverse passes then take exact dot4 distances
  const uint8_t* rows8;
  // segments: rows [seg_lo[c], seg_lo[c+1]) form independent data sets (null: one segment)
  const uint32_t* seg_of;  // [N] segment of each row
  const uint32_t* seg_lo;  // [nseg + 1]
  const uint32_t* perm;    // processing order of the rows (null: 0..N-1)
  int phase_clock;         // 1: k_forward adds its collect / filter cycles to counters[6], [7]
};

// MFMA screen of DPG's pair distances (see prune_screen.hip): the norms of the screen bound,
// the BF16 or FP8 copy, and the bound's relative slack
struct DpgScreen {
  const float* nrm2;
  const float* nq;
  const float* ne;
  const float* sc8;
  const uint16_t* Xb;
  const uint8_t* Xq8;
  float eta;
};
// precision: KP_BF16 or KP_FP8; K <= 64
void launch_dpg_screen(const PruneArgs& a, const DpgScreen& q, int precision, hipStream_t s);

// PH_EP runs one workgroup per segment: segment c starts its beam at its first row and
// aims at centroid + c * D, writing ep[c]
void launch_prune(const PruneArgs& a, int phase, const float* centroid, uint32_t* ep,
                  hipStream_t s, uint32_t nseg = 1);
void launch_centroid(const float* X, uint32_t N, uint32_t D, double* part, float* c,
                     hipStream_t s);

// ---------------------------------------------------------------------- kNN construction
enum KnnPrecision { KP_EXACT = 0, KP_FP32 = 1, KP_BF16 = 2, KP_FP8 = 3 };

struct KnnArgs {
  const float* X;
  uint32_t N, D, K;
  uint32_t* G;      // [N*K]
  float* GD;        // [N*K]
  uint8_t* V;       // [N*K] visited flags
  uint32_t* G1;     // phase-2 output ids
  float* GD1;       // phase-2 output distances
  uint32_t* H;      // [N*128] sample lists: new' at [0,64), old' at [64,128)
  uint32_t* na;     // [N] |new'|
  uint32_t* nb;     // [N] |old'|
  uint32_t* nnew;   // [N] |new| before reverse samples
  uint32_t* nold;   // [N]
  PushBuf pn, po, pc, pr;
  uint64_t* R0;     // [N*K] reverse lists (keys), read side
  uint32_t* Rn0;    // [N]
  uint64_t* R1;     // [N*K] write side
  uint32_t* Rn1;    // [N]
  uint64_t seed;
  unsigned long long* counters;  // [8] or null
  // screening (precision FP32 / BF16): a screen distance s = |a|^2 + |b|^2 - 2 a'.b' on the
  // screen copies a', b' with a rigorous bound |s - dist(a,b)| <= E(a,b); only candidates
  // whose bound interval can change a decision are re-scored with the canonical distance.
  const uint16_t* Xb;  // [N*D] BF16 rows (precision BF16)
  const float* nrm2;   // [N] |x|^2 (FP64 sum rounded once)
  const float* nq;     // [N] |x'| of the screen copy (rounded up)
  const float* ne;     // [N] |x' - x| (rounded up)
  float eta;           // relative slack of the bound
  // Integer data set (values in [-256, 256], |x|^2 small enough that every partial sum is
  // an integer below 2^24): the MFMA screen distance IS the canonical distance, no bound
  // and no re-score are needed.
  int exact_screen;
  // rows used for canonical re-scores: BF16 copy when the data set is BF16-exact, else X
  const uint16_t* Xr16;
  const uint8_t* X8;     // u8 copy of X (every value an integer in [0, 255], D <= 256) or null
  const float* n2s;      // [N] |x - 128|^2 of the u8 rows (exact), with X8
  // FP8 e4m3 copy with a per-vector scale (FP8 screen: nq/ne then hold its norms)
  const uint8_t* Xq8;
  const float* sc8;
  // segments (see PruneArgs): the initial lists draw ids inside the row's own segment, from a
  // stream of (segment seed, row index within the segment); every later step follows edges
  const uint32_t* seg_of;
  const uint32_t* seg_lo;
  const uint32_t* seg_key;  // [nseg] seed key of each segment
  const uint32_t* perm;     // processing order of the gather-heavy kernels (null: 0..N-1)
};

// FP8 copy: q = e4m3(x / s), s = max|x| / 440, with |s q| and |x - s q| (rounded up)
void launch_screen_prep8(const float* X, uint32_t N, uint32_t D, uint8_t* Xq, float* sc,
                         float* nq, float* ne, hipStream_t s);
// u8 copy of the n values of X (n % 4 == 0; every value an integer in [0, 255])
void launch_make_u8(const float* X, size_t n, uint8_t* X8, hipStream_t s);
// |x - 128|^2 of every u8 row (the norms of the rows shifted into int8)
void launch_u8_shift_norms(const uint8_t* X8, uint32_t N, uint32_t D, float* n2s, hipStream_t s);
// screen-mode preparation: BF16 copy (if Xb), norms, quantisation-error norms and the
// data-set certificates in flags[5] (zeroed by the caller)
void launch_screen_prep(const float* X, uint32_t N, uint32_t D, uint16_t* Xb, float* nrm2,
                        float* nq, float* ne, uint32_t* flags, hipStream_t s);

void launch_knn_init(const KnnArgs& a, hipStream_t s);
// canonical distances of the ids in G, rows sorted by (d, id), flags cleared
void launch_knn_init_sort(const KnnArgs& a, hipStream_t s);
void launch_knn_sample(const KnnArgs& a, hipStream_t s);
void launch_knn_rsample(const KnnArgs& a, hipStream_t s);
// tile: staging depth of the MFMA screen (0: 256-byte rows per round, 1: 64-byte rows)
void launch_knn_localjoin(const KnnArgs& a, int precision, int tile, hipStream_t s);
void launch_knn_merge(const KnnArgs& a, hipStream_t s);
// Reverse lists as exact runs per destination in a.pr (cnt, ov_off, ov_sorted; read them with
// cap 0).  hist holds rev_hist_words(N, K) words, scan_tmp scan_tmp_words(that - 1).
// Returns nonzero, launching nothing, beyond 2^26 ids or the edge capacity of a.pr.
size_t rev_hist_words(uint32_t N, uint32_t K);
int launch_knn_rev_build(const KnnArgs& a, uint32_t* hist, uint32_t* scan_tmp, hipStream_t s);
void launch_knn_rev_reduce(const KnnArgs& a, hipStream_t s);
void launch_knn_phase2(const KnnArgs& a, int precision, hipStream_t s);
void launch_knn_rupdate(const KnnArgs& a, hipStream_t s);
void launch_knn_final(const KnnArgs& a, hipStream_t s);

// ------------------------------------------------------------------------------ utilities
// push-buffer bookkeeping
void pb_reset(const PushBuf& pb, uint32_t N, uint32_t* fill, hipStream_t s);
void pb_group_overflow(const PushBuf& pb, uint32_t N, uint32_t* fill, uint32_t* scan_tmp,
                       hipStream_t s);
size_t scan_tmp_words(uint32_t N);
// exclusive scan of v[0..n) in place, v[n] = total
void excl_scan_u32(uint32_t* v, uint32_t n, uint32_t* tmp, hipStream_t s);

void fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);
// Beam search over a pruned graph (search.hip): queries Q [nq*D], pool L <= 256, k <= L;
// bitmap: `words` u32 per workgroup, all zero on entry and exit; vlist: vcap u32 per
// workgroup.  counters: [0] distance evaluations, [1] expansions.
struct SearchArgs {
  const float* X;
  uint32_t N, D;
  const uint32_t* deg;
  const uint32_t* ids;
  uint32_t FD, ep;
  const float* Q;
  uint32_t nq, L, k;
  uint64_t seed;
  uint32_t* bitmap;
  uint32_t words;
  uint32_t* vlist;
  uint32_t vcap;
  uint32_t* out_ids;
  float* out_d;
  unsigned long long* counters;
};
void launch_search(const SearchArgs& a, uint32_t grid, hipStream_t s);
void launch_selftest_sort(const uint64_t* in, uint64_t* out, int npl, int mode, uint32_t groups,
                          hipStream_t s);
void fill_u8(uint8_t* p, uint8_t v, size_t n, hipStream_t s);

// synthetic data: kind 0 integer GMM in [0,255], 1 unit-norm GMM, 2 |GMM| scaled to [0,1]
void launch_gen_data(int kind, uint32_t N, uint32_t D, uint32_t ncomp, uint64_t seed,
                     float sigma, float* centres, float* X, uint32_t* scratch, hipStream_t s);

// FP64 recall judge
void launch_recall(const float* X, uint32_t N, uint32_t D, const uint32_t* G, uint32_t K,
                   const uint32_t* q, uint32_t nq, uint32_t k, double* kth, uint32_t* cand_n,
                   double* cand_d, uint32_t cand_cap, uint32_t* hits, hipStream_t s,
                   const float* Q = nullptr);
// largest relative difference of the stored FP32 kNN distances of rows q[0..nq) from FP64
void launch_dist_check(const float* X, uint32_t N, uint32_t D, const uint32_t* G,
                       const float* GD, uint32_t K, const uint32_t* q, uint32_t nq,
                       unsigned long long* max_bits, hipStream_t s);

}  // namespace tg
