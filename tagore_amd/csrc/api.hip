// api.hip -- host orchestration and the C ABI declared in include/tagore.h.
//
// Owns every device buffer of a handle, drives the GNN-Descent schedule (pybind.cu:19-142)
// and the CFS dispatch (pybind.cu:145-295) on one HIP stream, times both with HIP events,
// and writes the index files (common.cu:151-248).  Errors are returned as status codes with
// a thread-local message; nothing here calls exit().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/tagore.h"
#include "errors.hpp"
#include "kernels.hpp"
#include "learn.hpp"

using namespace tg;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

}  // namespace

int tg::set_error(int code, const std::string& msg) { return fail(code, msg); }

namespace {

#define HIPX(expr)                                                                   \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fail(TG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
  } while (0)

template <class T>
int dalloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(p), n * sizeof(T));
  if (e != hipSuccess)
    return fail(TG_ERR_HIP, "hipMalloc of " + std::to_string(n * sizeof(T)) +
                                " bytes: " + hipGetErrorString(e));
  return TG_OK;
}
template <class T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

// The two window events of a build or a prune, destroyed on every return path.
struct EventPair {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  EventPair() = default;
  EventPair(const EventPair&) = delete;
  EventPair& operator=(const EventPair&) = delete;
  ~EventPair() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
};

struct PB {
  PushBuf d{};
  uint32_t N = 0;
  bool alloc(uint32_t n, uint32_t cap, size_t ov_cap) {
    N = n;
    d.cap = cap;
    d.ov_cap = uint32_t(std::min<size_t>(ov_cap, 0xFFFFFFF0u));
    if (dalloc(&d.cnt, n) || dalloc(&d.slots, size_t(n) * cap) || dalloc(&d.ov_dest, d.ov_cap) ||
        dalloc(&d.ov_key, d.ov_cap) || dalloc(&d.ov_used, 1) || dalloc(&d.ov_off, size_t(n) + 1) ||
        dalloc(&d.ov_sorted, d.ov_cap) || dalloc(&d.err, 1))
      return false;
    return true;
  }
  void release() {
    dfree(d.cnt);
    dfree(d.slots);
    dfree(d.ov_dest);
    dfree(d.ov_key);
    dfree(d.ov_used);
    dfree(d.ov_off);
    dfree(d.ov_sorted);
    dfree(d.err);
  }
};

}  // namespace

struct tg_state {
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t N = 0, D = 0;
  float* X = nullptr;
  float* centres = nullptr;  // synthetic-data centres
  uint32_t ncomp_alloc = 0;
  uint32_t* misc = nullptr;  // [64] small scratch
  // kNN
  uint32_t K = 0;
  bool has_knn = false;
  uint32_t* G = nullptr;
  float* GD = nullptr;
  uint8_t* V = nullptr;
  uint32_t K_scratch = 0;
  uint32_t *G1 = nullptr, *H = nullptr, *na = nullptr, *nb = nullptr, *nnew = nullptr,
           *nold = nullptr;
  float* GD1 = nullptr;
  uint64_t *R0 = nullptr, *R1 = nullptr;
  uint32_t *Rn0 = nullptr, *Rn1 = nullptr;
  PB pn, po, pc, pr;
  uint32_t *rv_hist = nullptr, *rv_scan = nullptr;  // reverse-run histogram and its scan
  uint8_t* Xq8 = nullptr;  // FP8 e4m3 copy of X with per-vector scales (screen)
  float *sc8 = nullptr, *nq8 = nullptr, *ne8 = nullptr;
  bool prep8_valid = false;
  uint16_t* Xb = nullptr;  // BF16 copy of X (screen)
  float *nrm2 = nullptr, *nq = nullptr, *ne = nullptr;
  uint32_t* cert_flags = nullptr;  // [5], see k_screen_prep
  bool prep_valid = false;         // Xb / norms / certificates match the current X
  bool cert_bf16 = false;          // every value of X is exactly representable in BF16
  bool cert_int = false;           // integer data whose screen distances are exact
  bool cert_u8 = false;            // and every value in [0, 255]: X8 holds the u8 copy
  uint8_t* X8 = nullptr;
  float* n2s = nullptr;            // [N] |x - 128|^2 of the u8 rows
  uint32_t* fill = nullptr;
  uint32_t* scan_tmp = nullptr;
  unsigned long long* counters = nullptr;  // [CNT_SLOTS * CNT_N], summed over the slots
  // prune
  uint32_t FD_alloc = 0;
  uint32_t *fwd_deg = nullptr, *fwd_ids = nullptr, *out_deg = nullptr, *out_ids = nullptr;
  float* fwd_d = nullptr;
  float* centroid = nullptr;
  double* cpart = nullptr;
  uint32_t* ep_dev = nullptr;
  PB rev;
  bool has_prune = false;
  tg_prune_params last_prune{};
  uint32_t ep = 0;
  // search validator: per-workgroup visited bitmaps (all zero between calls) and id lists
  uint32_t* sbm = nullptr;
  uint32_t* svl = nullptr;
  unsigned long long* scnt = nullptr;
  // segments (clusters stored back to back; nseg == 0: one data set)
  uint32_t nseg = 0;
  std::vector<uint32_t> seg_lo_host;
  uint32_t* seg_of = nullptr;   // [N]
  uint32_t* seg_lo = nullptr;   // [nseg + 1]
  uint32_t* seg_ep = nullptr;   // [nseg]
  uint32_t* seg_key = nullptr;  // [nseg] seed keys
  float* seg_cent = nullptr;    // [nseg * D]
  uint32_t seg_cap = 0;
  std::vector<uint32_t> seg_ep_host;
  // learned initial lists
  Learner learn;
  tg_state* sub = nullptr;     // the sample's own handle (its kNN graph trains the encoder)
  uint32_t* samp = nullptr;    // sample ids
  uint32_t samp_cap = 0;
  bool has_codes = false;
  bool order_valid = false;    // learn.order (bucket order) belongs to the current graph
  // stats
  int profiling = 0;  // 1: per-kernel events, 2: also prune phase clocks
  tg_stats stats{};
  std::vector<hipEvent_t> ev_pool;
  // captured iteration sequences of tg_knn_build (TG_FLAG_GRAPH_ITERS), keyed by their
  // launch arguments; two entries cover the G / G1 alternation of odd expansion counts
  std::vector<std::pair<std::vector<unsigned char>, hipGraphExec_t>> iter_graphs;
};

namespace {

int ensure_events(tg_state* s, size_t n) {
  while (s->ev_pool.size() < n) {
    hipEvent_t e;
    HIPX(hipEventCreate(&e));
    s->ev_pool.push_back(e);
  }
  return TG_OK;
}

int check_pb_err(tg_state* s, PB& pb, const char* what) {
  uint32_t e = 0;
  HIPX(hipMemcpyAsync(&e, pb.d.err, 4, hipMemcpyDeviceToHost, s->stream));
  HIPX(hipStreamSynchronize(s->stream));
  if (e)
    return fail(TG_ERR_CAPACITY, std::string("push-buffer overflow area exhausted in ") + what);
  return TG_OK;
}

int alloc_pb(tg_state* s, PB& pb, uint32_t cap, size_t ov_cap) {
  if (pb.d.cnt && pb.d.cap == cap && pb.d.ov_cap >= std::min<size_t>(ov_cap, 0xFFFFFFF0u) &&
      pb.N == s->N)
    return TG_OK;
  pb.release();
  if (!pb.alloc(s->N, cap, ov_cap)) return TG_ERR_HIP;
  fill_u32(pb.d.err, 0, 1, s->stream);
  return TG_OK;
}

void clear_errors(tg_state* s, std::initializer_list<PB*> pbs) {
  for (PB* pb : pbs)
    if (pb->d.err) fill_u32(pb->d.err, 0, 1, s->stream);
}

int ensure_common_scratch(tg_state* s) {
  if (!s->fill) {
    if (dalloc(&s->fill, s->N) || dalloc(&s->scan_tmp, scan_tmp_words(s->N) + 16) ||
        dalloc(&s->counters, size_t(CNT_SLOTS) * CNT_N))
      return TG_ERR_HIP;
  }
  return TG_OK;
}

int ensure_knn_scratch(tg_state* s, uint32_t K) {
  const size_t N = s->N;
  if (s->K_scratch != K) {
    dfree(s->G1);
    dfree(s->GD1);
    dfree(s->R0);
    dfree(s->R1);
    dfree(s->Rn0);
    dfree(s->Rn1);
    dfree(s->rv_hist);
    dfree(s->rv_scan);
    const size_t hw = rev_hist_words(uint32_t(N), K);
    if (dalloc(&s->G1, N * K) || dalloc(&s->GD1, N * K) || dalloc(&s->R0, N * K) ||
        dalloc(&s->R1, N * K) || dalloc(&s->Rn0, N) || dalloc(&s->Rn1, N) ||
        dalloc(&s->rv_hist, hw) || dalloc(&s->rv_scan, scan_tmp_words(uint32_t(hw - 1)) + 16))
      return TG_ERR_HIP;
    s->K_scratch = K;
  }
  if (!s->H) {
    if (dalloc(&s->H, N * 128) || dalloc(&s->na, N) || dalloc(&s->nb, N) ||
        dalloc(&s->nnew, N) || dalloc(&s->nold, N))
      return TG_ERR_HIP;
  }
  if (ensure_common_scratch(s)) return TG_ERR_HIP;
  if (alloc_pb(s, s->pn, 64, N * 32) || alloc_pb(s, s->po, 64, N * 32) ||
      alloc_pb(s, s->pc, 256, N * 256) || alloc_pb(s, s->pr, 128, N * size_t(K)))
    return TG_ERR_HIP;
  return TG_OK;
}

// BF16 copy, norms and data-set certificates of the current vectors (computed once per data
// set, like the reference's host-side FP16 copy, pybind.cu:58-61).
int ensure_prep(tg_state* s) {
  if (s->prep_valid) return TG_OK;
  const size_t N = s->N;
  if (!s->nrm2) {
    if (dalloc(&s->nrm2, N) || dalloc(&s->nq, N) || dalloc(&s->ne, N) ||
        dalloc(&s->cert_flags, 5) || dalloc(&s->Xb, N * s->D))
      return TG_ERR_HIP;
  }
  fill_u32(s->cert_flags, 0, 5, s->stream);
  launch_screen_prep(s->X, s->N, s->D, s->Xb, s->nrm2, s->nq, s->ne, s->cert_flags, s->stream);
  uint32_t f[5];
  HIPX(hipMemcpyAsync(f, s->cert_flags, sizeof(f), hipMemcpyDeviceToHost, s->stream));
  HIPX(hipStreamSynchronize(s->stream));
  s->cert_bf16 = f[0] == 0;
  float qmax;
  std::memcpy(&qmax, &f[3], 4);
  // every MFMA partial sum, |a|^2+|b|^2 and |a-b|^2 stay integers below 2^24
  const double lim = 16777216.0;
  s->cert_int = f[1] == 0 && (f[2] == 0 ? 2.0 * qmax < lim : 4.0 * qmax < lim);
  // non-negative integers up to 255 with every |a-b|^2 below 2^24: the exact u8 dot4
  // distance equals the canonical FP32 fold
  s->cert_u8 = s->cert_int && f[2] == 0 && f[4] == 0 && s->D % 4 == 0;
  if (s->cert_u8) {
    if (!s->X8 && dalloc(&s->X8, N * s->D)) return TG_ERR_HIP;
    if (!s->n2s && dalloc(&s->n2s, N)) return TG_ERR_HIP;
    launch_make_u8(s->X, N * s->D, s->X8, s->stream);
    if (s->D <= 256) launch_u8_shift_norms(s->X8, s->N, s->D, s->n2s, s->stream);
  }
  s->prep_valid = true;
  return TG_OK;
}

// FP8 copy with per-vector scales and its norms (once per data set)
int ensure_prep8(tg_state* s) {
  if (s->prep8_valid) return TG_OK;
  const size_t N = s->N;
  if (!s->Xq8) {
    if (dalloc(&s->Xq8, N * s->D) || dalloc(&s->sc8, N) || dalloc(&s->nq8, N) ||
        dalloc(&s->ne8, N))
      return TG_ERR_HIP;
  }
  launch_screen_prep8(s->X, s->N, s->D, s->Xq8, s->sc8, s->nq8, s->ne8, s->stream);
  s->prep8_valid = true;
  return TG_OK;
}

int ensure_graph(tg_state* s, uint32_t K) {
  if (s->K != K || !s->G) {
    dfree(s->G);
    dfree(s->GD);
    dfree(s->V);
    if (dalloc(&s->G, size_t(s->N) * K) || dalloc(&s->GD, size_t(s->N) * K) ||
        dalloc(&s->V, size_t(s->N) * K))
      return TG_ERR_HIP;
    s->K = K;
  }
  return TG_OK;
}

KnnArgs knn_args(tg_state* s, uint64_t seed) {
  KnnArgs a{};
  a.X = s->X;
  a.N = s->N;
  a.D = s->D;
  a.K = s->K;
  a.G = s->G;
  a.GD = s->GD;
  a.V = s->V;
  a.G1 = s->G1;
  a.GD1 = s->GD1;
  a.H = s->H;
  a.na = s->na;
  a.nb = s->nb;
  a.nnew = s->nnew;
  a.nold = s->nold;
  a.pn = s->pn.d;
  a.po = s->po.d;
  a.pc = s->pc.d;
  a.pr = s->pr.d;
  a.R0 = s->R0;
  a.Rn0 = s->Rn0;
  a.R1 = s->R1;
  a.Rn1 = s->Rn1;
  a.seed = seed;
  a.counters = s->counters;
  a.Xb = s->Xb;
  a.nrm2 = s->nrm2;
  a.nq = s->nq;
  a.ne = s->ne;
  // relative slack of the screen bound: accumulation (2 D ulps per term, doubled) plus the
  // canonical distance's own rounding, with margin
  a.eta = float((8.0 * s->D + 128.0) * 5.9604644775390625e-8);
  return a;
}

// event-timed section helper
struct Sections {
  tg_state* s;
  std::vector<std::pair<int, size_t>> rec;  // (slot, event index of start)
  size_t next = 0;
  int start(int slot) {
    if (!s->profiling) return TG_OK;
    if (ensure_events(s, next + 2)) return TG_ERR_HIP;
    HIPX(hipEventRecord(s->ev_pool[next], s->stream));
    rec.push_back({slot, next});
    next += 2;
    return TG_OK;
  }
  int stop() {
    if (!s->profiling) return TG_OK;
    HIPX(hipEventRecord(s->ev_pool[rec.back().second + 1], s->stream));
    return TG_OK;
  }
  int collect(double* out) {
    for (auto& r : rec) {
      float ms = 0.f;
      HIPX(hipEventElapsedTime(&ms, s->ev_pool[r.second], s->ev_pool[r.second + 1]));
      out[r.first] += ms;
    }
    return TG_OK;
  }
};

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(TG_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return TG_OK;
}

// the evaluation counters, summed over their slots
int read_counters(tg_state* s, unsigned long long (&cnt)[CNT_N]) {
  std::vector<unsigned long long> all(size_t(CNT_SLOTS) * CNT_N);
  HIPX(hipMemcpy(all.data(), s->counters, all.size() * 8, hipMemcpyDeviceToHost));
  for (int i = 0; i < CNT_N; ++i) {
    cnt[i] = 0;
    for (int b = 0; b < CNT_SLOTS; ++b) cnt[i] += all[size_t(b) * CNT_N + i];
  }
  return TG_OK;
}

uint64_t splitmix64(uint64_t& x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

}  // namespace

// ===================================================================== C ABI
// Host-side exceptions (allocation failures of data-sized host buffers, mostly) become
// status codes: nothing unwinds through the C boundary.
template <class F>
static int guarded(F&& f) {
  try {
    return f();
  } catch (const std::bad_alloc&) {
    return fail(TG_ERR_CAPACITY, "host memory exhausted");
  } catch (const std::exception& e) {
    return fail(TG_ERR_STATE, e.what());
  }
}

extern "C" {

const char* tg_last_error(void) { return g_err.c_str(); }

int tg_create(uint32_t N, uint32_t D, int device, tg_state** out) {
  if (!out) return fail(TG_ERR_ARG, "out is null");
  *out = nullptr;
  if (N < 2) return fail(TG_ERR_ARG, "N must be >= 2");
  if (D == 0 || D % 4 != 0 || D > 1024)
    return fail(TG_ERR_ARG, "D must be a multiple of 4 in [4, 1024]");
  int ndev = 0;
  HIPX(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(TG_ERR_ARG, "no such HIP device");
  HIPX(hipSetDevice(device));
  tg_state* s = new tg_state();
  s->device = device;
  s->N = N;
  s->D = D;
  hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete s;
    return fail(TG_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  if (dalloc(&s->X, size_t(N) * D) || dalloc(&s->misc, 64) || dalloc(&s->centroid, D) ||
      dalloc(&s->ep_dev, 1)) {
    std::string m = g_err;
    tg_free(s);
    return fail(TG_ERR_HIP, m);
  }
  *out = s;
  return TG_OK;
}

int tg_free(tg_state* s) {
  if (!s) return TG_OK;
  (void)hipSetDevice(s->device);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  tg_free(s->sub);
  s->sub = nullptr;
  for (auto& g : s->iter_graphs) (void)hipGraphExecDestroy(g.second);
  s->iter_graphs.clear();
  learner_release(s->learn);
  dfree(s->samp);
  dfree(s->seg_of);
  dfree(s->seg_lo);
  dfree(s->seg_ep);
  dfree(s->seg_key);
  dfree(s->seg_cent);
  dfree(s->X);
  dfree(s->centres);
  dfree(s->misc);
  dfree(s->G);
  dfree(s->GD);
  dfree(s->V);
  dfree(s->G1);
  dfree(s->GD1);
  dfree(s->H);
  dfree(s->na);
  dfree(s->nb);
  dfree(s->nnew);
  dfree(s->nold);
  dfree(s->R0);
  dfree(s->R1);
  dfree(s->Rn0);
  dfree(s->Rn1);
  dfree(s->rv_hist);
  dfree(s->rv_scan);
  s->pn.release();
  s->po.release();
  s->pc.release();
  s->pr.release();
  s->rev.release();
  dfree(s->Xb);
  dfree(s->Xq8);
  dfree(s->sc8);
  dfree(s->nq8);
  dfree(s->ne8);
  dfree(s->nrm2);
  dfree(s->nq);
  dfree(s->ne);
  dfree(s->cert_flags);
  dfree(s->X8);
  dfree(s->n2s);
  dfree(s->fill);
  dfree(s->scan_tmp);
  dfree(s->counters);
  dfree(s->sbm);
  dfree(s->svl);
  dfree(s->scnt);
  dfree(s->fwd_deg);
  dfree(s->fwd_ids);
  dfree(s->fwd_d);
  dfree(s->out_deg);
  dfree(s->out_ids);
  dfree(s->centroid);
  dfree(s->cpart);
  dfree(s->ep_dev);
  for (auto e : s->ev_pool) (void)hipEventDestroy(e);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return TG_OK;
}

int tg_set_data(tg_state* s, const float* rows) {
  if (!s || !rows) return fail(TG_ERR_ARG, "null argument");
  HIPX(hipSetDevice(s->device));
  HIPX(hipMemcpyAsync(s->X, rows, size_t(s->N) * s->D * 4, hipMemcpyHostToDevice, s->stream));
  HIPX(hipStreamSynchronize(s->stream));
  s->has_knn = false;
  s->prep_valid = false;
  s->prep8_valid = false;
  s->has_prune = false;
  s->order_valid = false;
  return TG_OK;
}

int tg_gen_data(tg_state* s, int32_t kind, uint32_t ncomp, uint64_t seed, float sigma) {
  if (!s) return fail(TG_ERR_ARG, "null handle");
  if (kind < 0 || kind > 2 || ncomp == 0) return fail(TG_ERR_ARG, "bad data kind / ncomp");
  HIPX(hipSetDevice(s->device));
  if (s->ncomp_alloc < ncomp) {
    dfree(s->centres);
    if (dalloc(&s->centres, size_t(ncomp) * s->D)) return TG_ERR_HIP;
    s->ncomp_alloc = ncomp;
  }
  launch_gen_data(kind, s->N, s->D, ncomp, seed, sigma, s->centres, s->X, s->misc, s->stream);
  if (int rc = check_launch("gen_data")) return rc;
  HIPX(hipStreamSynchronize(s->stream));
  s->has_knn = false;
  s->prep_valid = false;
  s->prep8_valid = false;
  s->has_prune = false;
  s->order_valid = false;
  return TG_OK;
}

int tg_get_data(tg_state* s, float* rows, uint32_t first, uint32_t count) {
  if (!s || !rows) return fail(TG_ERR_ARG, "null argument");
  if (uint64_t(first) + count > s->N) return fail(TG_ERR_ARG, "range outside the data");
  HIPX(hipSetDevice(s->device));
  HIPX(hipMemcpy(rows, s->X + size_t(first) * s->D, size_t(count) * s->D * 4,
                 hipMemcpyDeviceToHost));
  return TG_OK;
}

int tg_read_fvecs(const char* path, float** rows, uint32_t* N, uint32_t* D) {
  if (!path || !rows || !N || !D) return fail(TG_ERR_ARG, "null argument");
  *rows = nullptr;
  FILE* f = fopen(path, "rb");
  if (!f) return fail(TG_ERR_IO, std::string("cannot open ") + path);
  uint32_t dim = 0;
  if (fread(&dim, 4, 1, f) != 1 || dim == 0) {
    fclose(f);
    return fail(TG_ERR_IO, std::string("short or empty fvecs file ") + path);
  }
  fseek(f, 0, SEEK_END);
  const long long size = ftell(f);
  const long long rec = 4ll * (dim + 1);
  if (size % rec != 0) {
    fclose(f);
    return fail(TG_ERR_IO, std::string("fvecs size is not a multiple of the record: ") + path);
  }
  const uint64_t n = uint64_t(size / rec);
  float* out = static_cast<float*>(malloc(n * dim * 4 + 4));
  if (!out) {
    fclose(f);
    return fail(TG_ERR_IO, "out of host memory");
  }
  fseek(f, 0, SEEK_SET);
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t d2 = 0;
    if (fread(&d2, 4, 1, f) != 1 || d2 != dim || fread(out + i * dim, 4, dim, f) != dim) {
      fclose(f);
      free(out);
      return fail(TG_ERR_IO, std::string("inconsistent fvecs record in ") + path);
    }
  }
  fclose(f);
  *rows = out;
  *N = uint32_t(n);
  *D = dim;
  return TG_OK;
}

// Host readers of the other vector formats.  bvecs: per vector u32 dim then dim bytes;
// fbin: u32 N, u32 D, then N*D f32 (the big-ann layout); both give f32 rows.
int tg_read_bvecs(const char* path, float** rows, uint32_t* N, uint32_t* D) {
  if (!path || !rows || !N || !D) return fail(TG_ERR_ARG, "null argument");
  *rows = nullptr;
  FILE* f = fopen(path, "rb");
  if (!f) return fail(TG_ERR_IO, std::string("cannot open ") + path);
  uint32_t dim = 0;
  if (fread(&dim, 4, 1, f) != 1 || dim == 0) {
    fclose(f);
    return fail(TG_ERR_IO, std::string("short or empty bvecs file ") + path);
  }
  fseek(f, 0, SEEK_END);
  const long long size = ftell(f);
  const long long rec = 4ll + dim;
  if (size % rec != 0) {
    fclose(f);
    return fail(TG_ERR_IO, std::string("bvecs size is not a multiple of the record: ") + path);
  }
  const uint64_t n = uint64_t(size / rec);
  float* out = static_cast<float*>(malloc(n * dim * 4 + 4));
  std::vector<uint8_t> buf(dim);
  if (!out) {
    fclose(f);
    return fail(TG_ERR_IO, "out of host memory");
  }
  fseek(f, 0, SEEK_SET);
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t d2 = 0;
    if (fread(&d2, 4, 1, f) != 1 || d2 != dim || fread(buf.data(), 1, dim, f) != dim) {
      fclose(f);
      free(out);
      return fail(TG_ERR_IO, std::string("inconsistent bvecs record in ") + path);
    }
    for (uint32_t j = 0; j < dim; ++j) out[i * dim + j] = float(buf[j]);
  }
  fclose(f);
  *rows = out;
  *N = uint32_t(n);
  *D = dim;
  return TG_OK;
}

int tg_read_fbin(const char* path, float** rows, uint32_t* N, uint32_t* D) {
  if (!path || !rows || !N || !D) return fail(TG_ERR_ARG, "null argument");
  *rows = nullptr;
  FILE* f = fopen(path, "rb");
  if (!f) return fail(TG_ERR_IO, std::string("cannot open ") + path);
  uint32_t hd[2] = {0, 0};
  if (fread(hd, 4, 2, f) != 2 || hd[0] == 0 || hd[1] == 0) {
    fclose(f);
    return fail(TG_ERR_IO, std::string("short or empty fbin header in ") + path);
  }
  fseek(f, 0, SEEK_END);
  const long long size = ftell(f);
  const uint64_t want = 8ull + 4ull * uint64_t(hd[0]) * hd[1];
  if (uint64_t(size) != want) {
    fclose(f);
    return fail(TG_ERR_IO, std::string("fbin size does not match its header: ") + path);
  }
  const uint64_t n = uint64_t(hd[0]) * hd[1];
  float* out = static_cast<float*>(malloc(n * 4 + 4));
  if (!out) {
    fclose(f);
    return fail(TG_ERR_IO, "out of host memory");
  }
  fseek(f, 8, SEEK_SET);
  if (fread(out, 4, n, f) != n) {
    fclose(f);
    free(out);
    return fail(TG_ERR_IO, std::string("short fbin file ") + path);
  }
  fclose(f);
  *rows = out;
  *N = hd[0];
  *D = hd[1];
  return TG_OK;
}

// Reader of the five index layouts tg_store_index writes (the reference's StoreNSG ...
// StoreCAGRA, common.cu:151-248).  N rows; width = the row capacity (NSG / NSSG header, the
// Vamana max degree, the DPG largest degree, the CAGRA d); deg [N] and ids [N*width]
// (malloc'ed, padded with INV); eps: the entry points the file names (NSG / Vamana: 1,
// NSSG: 10, DPG / CAGRA: 0).
static int read_index_impl(const char* path, const char* index_type, uint32_t N,
                           uint32_t* width, uint32_t** deg, uint32_t** ids, uint32_t* eps,
                           uint32_t* n_eps) {
  if (!path || !index_type || !width || !deg || !ids || !eps || !n_eps)
    return fail(TG_ERR_ARG, "null argument");
  *deg = nullptr;
  *ids = nullptr;
  *n_eps = 0;
  const std::string t(index_type);
  int kind = -1;
  if (t == "NSG") kind = 0;
  else if (t == "Vamana") kind = 1;
  else if (t == "NSSG") kind = 2;
  else if (t == "DPG") kind = 3;
  else if (t == "CAGRA") kind = 4;
  else return fail(TG_ERR_ARG, "unknown index type " + t);
  if (N == 0) return fail(TG_ERR_ARG, "N must be >= 1");
  FILE* f = fopen(path, "rb");
  if (!f) return fail(TG_ERR_IO, std::string("cannot open ") + path);
  std::vector<uint32_t> words;
  {
    fseek(f, 0, SEEK_END);
    const long long size = ftell(f);
    fseek(f, 0, SEEK_SET);
    if (size < 0 || size % 4 != 0) {
      fclose(f);
      return fail(TG_ERR_IO, std::string("index size is not a multiple of 4: ") + path);
    }
    words.resize(size_t(size / 4));
    if (!words.empty() && fread(words.data(), 4, words.size(), f) != words.size()) {
      fclose(f);
      return fail(TG_ERR_IO, std::string("short read of ") + path);
    }
  }
  fclose(f);
  size_t pos = 0;
  auto bad = [&](const char* what) {
    return fail(TG_ERR_IO, std::string("malformed ") + t + " index (" + what + "): " + path);
  };
  auto take = [&](uint32_t& v) {
    if (pos >= words.size()) return false;
    v = words[pos++];
    return true;
  };
  uint32_t w = 0;
  if (kind == 0 || kind == 2) {
    if (!take(w)) return bad("header");
    if (kind == 0) {
      if (!take(eps[0])) return bad("header");
      *n_eps = 1;
    } else {
      uint32_t ne = 0;
      if (!take(ne) || ne > 10) return bad("entry points");
      for (uint32_t i = 0; i < ne; ++i)
        if (!take(eps[i])) return bad("entry points");
      *n_eps = ne;
    }
  } else if (kind == 1) {
    uint32_t lo = 0, hi = 0, f0 = 0, f1 = 0;
    if (!take(lo) || !take(hi) || !take(w) || !take(eps[0]) || !take(f0) || !take(f1))
      return bad("header");
    if ((uint64_t(hi) << 32 | lo) != uint64_t(words.size()) * 4) return bad("index_size");
    *n_eps = 1;
  } else if (kind == 3) {
    uint32_t m0 = 0, m1 = 0, vmaj = 0, vmin = 0, n = 0;
    if (!take(m0) || !take(m1) || !take(vmaj) || !take(vmin) || !take(n)) return bad("header");
    if (std::memcmp(&m0, "KNNG", 4) != 0 || std::memcmp(&m1, "RAPH", 4) != 0 || n != N)
      return bad("magic or N");
  }
  // rows: kind 3 repeats the degree; kind 4 rows have no separate degree (width d each)
  const size_t body = pos;
  if (kind == 3 || kind == 4) {
    w = 0;
    for (uint32_t u = 0; u < N; ++u) {
      uint32_t d = 0;
      if (!take(d)) return bad("rows");
      if (kind == 3) {
        uint32_t d2 = 0;
        if (!take(d2) || d2 != d) return bad("degree pair");
      }
      if (pos + d > words.size()) return bad("rows");
      pos += d;
      w = std::max(w, d);
    }
    pos = body;
  }
  if (w == 0 || w > 128) return bad("row width must be in [1, 128]");
  uint32_t* dg = static_cast<uint32_t*>(malloc(size_t(N) * 4));
  uint32_t* id = static_cast<uint32_t*>(malloc(size_t(N) * w * 4));
  if (!dg || !id) {
    free(dg);
    free(id);
    return fail(TG_ERR_IO, "out of host memory");
  }
  int rc = TG_OK;
  for (uint32_t u = 0; u < N && rc == TG_OK; ++u) {
    uint32_t d = 0;
    if (!take(d)) rc = bad("rows");
    if (rc == TG_OK && kind == 3) pos += 1;
    if (rc == TG_OK && (d > w || pos + d > words.size())) rc = bad("row degree");
    if (rc != TG_OK) break;
    dg[u] = d;
    for (uint32_t j = 0; j < w; ++j) {
      const uint32_t v = j < d ? words[pos + j] : TG_INV;
      if (j < d && v >= N) rc = bad("id out of range");
      id[size_t(u) * w + j] = v;
    }
    pos += d;
  }
  if (rc == TG_OK && pos != words.size()) rc = bad("trailing data or wrong N");
  for (uint32_t i = 0; rc == TG_OK && i < *n_eps; ++i)
    if (eps[i] >= N) rc = bad("entry point out of range");
  if (rc != TG_OK) {
    free(dg);
    free(id);
    return rc;
  }
  *width = w;
  *deg = dg;
  *ids = id;
  return TG_OK;
}
int tg_read_index(const char* path, const char* index_type, uint32_t N, uint32_t* width,
                  uint32_t** deg, uint32_t** ids, uint32_t* eps, uint32_t* n_eps) {
  return guarded(
      [&] { return read_index_impl(path, index_type, N, width, deg, ids, eps, n_eps); });
}

// Installs a pruned graph (deg [N], ids [N*width], entry point) as the handle's current one,
// e.g. an index read with tg_read_index, so that tg_search and tg_store_index use it.
int tg_set_pruned(tg_state* s, uint32_t width, const uint32_t* deg, const uint32_t* ids,
                  uint32_t ep) {
  if (!s || !deg || !ids) return fail(TG_ERR_ARG, "null argument");
  if (s->nseg) return fail(TG_ERR_STATE, "a pruned graph is set on a handle without segments");
  if (width < 1 || width > 128) return fail(TG_ERR_ARG, "width must be in [1, 128]");
  if (ep >= s->N) return fail(TG_ERR_ARG, "entry point out of range");
  for (uint32_t u = 0; u < s->N; ++u) {
    if (deg[u] > width) return fail(TG_ERR_ARG, "degree above width");
    for (uint32_t j = 0; j < deg[u]; ++j)
      if (ids[size_t(u) * width + j] >= s->N) return fail(TG_ERR_ARG, "id out of range");
  }
  HIPX(hipSetDevice(s->device));
  const uint32_t N = s->N;
  if (s->FD_alloc < width) {
    dfree(s->fwd_deg);
    dfree(s->fwd_ids);
    dfree(s->fwd_d);
    dfree(s->out_deg);
    dfree(s->out_ids);
    if (dalloc(&s->fwd_deg, N) || dalloc(&s->fwd_ids, size_t(N) * width) ||
        dalloc(&s->fwd_d, size_t(N) * width) || dalloc(&s->out_deg, N) ||
        dalloc(&s->out_ids, size_t(N) * width))
      return TG_ERR_HIP;
    s->FD_alloc = width;
  }
  HIPX(hipMemcpy(s->out_deg, deg, size_t(N) * 4, hipMemcpyHostToDevice));
  HIPX(hipMemcpy(s->out_ids, ids, size_t(N) * width * 4, hipMemcpyHostToDevice));
  s->last_prune = tg_prune_params{};
  s->last_prune.FD = width;
  s->ep = ep;
  s->has_prune = true;
  return TG_OK;
}

void tg_host_free(void* p) { free(p); }

int tg_set_profiling(tg_state* s, int on) {
  if (!s) return fail(TG_ERR_ARG, "null handle");
  s->profiling = on < 0 ? 0 : (on > 2 ? 2 : on);
  return TG_OK;
}

int tg_get_codes(tg_state* s, float* codes) {
  if (!s || !codes) return fail(TG_ERR_ARG, "null argument");
  if (!s->has_codes) return fail(TG_ERR_STATE, "no learned initialisation has run");
  HIPX(hipSetDevice(s->device));
  HIPX(hipMemcpy(codes, s->learn.codes, size_t(s->N) * MLP_CODE * 4, hipMemcpyDeviceToHost));
  return TG_OK;
}

namespace {

// device resources of the encoder checks, released on every exit path
struct MlpScratch {
  hipStream_t st = nullptr;
  Learner L;
  float* X = nullptr;
  float* g = nullptr;
  ~MlpScratch() {
    if (st) (void)hipStreamSynchronize(st);
    learner_release(L);
    dfree(X);
    dfree(g);
    if (st) (void)hipStreamDestroy(st);
  }
};

}  // namespace

int tg_selftest_sort(int device, uint32_t npl, uint32_t groups, int32_t mode,
                     const uint64_t* keys, uint64_t* out) {
  if (!keys || !out) return fail(TG_ERR_ARG, "null argument");
  if (npl != 1 && npl != 2 && npl != 8) return fail(TG_ERR_ARG, "npl must be 1, 2 or 8");
  if (mode < 0 || mode > 2 || (mode == 1 && npl != 1) || (mode == 2 && npl != 2))
    return fail(TG_ERR_ARG, "mode 0, mode 1 with npl 1, or mode 2 with npl 2");
  if (groups == 0 || groups > 65536) return fail(TG_ERR_ARG, "groups must be in [1, 65536]");
  HIPX(hipSetDevice(device));
  const size_t n = size_t(groups) * 64 * npl;
  uint64_t *din = nullptr, *dout = nullptr;
  if (dalloc(&din, n) || dalloc(&dout, n)) {
    dfree(din);
    return TG_ERR_HIP;
  }
  hipStream_t st;
  HIPX(hipStreamCreate(&st));
  HIPX(hipMemcpyAsync(din, keys, n * 8, hipMemcpyHostToDevice, st));
  launch_selftest_sort(din, dout, int(npl), mode, groups, st);
  int rc = check_launch("sort self-test");
  if (!rc) {
    HIPX(hipMemcpyAsync(out, dout, n * 8, hipMemcpyDeviceToHost, st));
    HIPX(hipStreamSynchronize(st));
  }
  HIPX(hipStreamDestroy(st));
  dfree(din);
  dfree(dout);
  return rc;
}

int tg_mlp_eval(int device, uint32_t D, uint32_t B, uint32_t P, const float* X, const float* mu,
                float scale, const float* params, float margin, float* loss, float* grad,
                float* codes) {
  if (!X || !mu || !params || !loss || !grad || !codes) return fail(TG_ERR_ARG, "null argument");
  if (D == 0 || D % 32 != 0 || D > 1024) return fail(TG_ERR_ARG, "D must be a multiple of 32 <= 1024");
  if (B < 32 || B % 32 != 0 || B > 4096) return fail(TG_ERR_ARG, "B must be a multiple of 32 in [32, 4096]");
  if (P < 1 || P > 16) return fail(TG_ERR_ARG, "P must be in [1, 16]");
  HIPX(hipSetDevice(device));
  const uint32_t M = B * (1 + P);
  const size_t n = size_t(mlp_n1(D)) + mlp_n2();
  MlpScratch w;
  HIPX(hipStreamCreate(&w.st));
  HIPX(learner_reserve(w.L, D, M, M, 1, 1));
  if (dalloc(&w.X, size_t(M) * D) || dalloc(&w.g, n)) return TG_ERR_HIP;
  HIPX(hipMemcpyAsync(w.X, X, size_t(M) * D * 4, hipMemcpyHostToDevice, w.st));
  HIPX(hipMemcpyAsync(w.L.mu, mu, size_t(D) * 4, hipMemcpyHostToDevice, w.st));
  HIPX(hipMemcpyAsync(w.L.scal, &scale, 4, hipMemcpyHostToDevice, w.st));
  HIPX(mlp_set_params(w.L, params, w.st));
  mlp_forward_backward(w.L, w.X, nullptr, B, P, margin, w.st);
  LearnConfig c;
  launch_mlp_adam(w.L, 1, 0, c, w.g, w.st);
  if (int rc = check_launch("encoder check")) return rc;
  std::vector<float> lp(B);
  HIPX(hipMemcpyAsync(lp.data(), w.L.lossp, size_t(B) * 4, hipMemcpyDeviceToHost, w.st));
  HIPX(hipMemcpyAsync(grad, w.g, n * 4, hipMemcpyDeviceToHost, w.st));
  HIPX(hipMemcpyAsync(codes, w.L.Z, size_t(M) * MLP_CODE * 4, hipMemcpyDeviceToHost, w.st));
  HIPX(hipStreamSynchronize(w.st));
  double sum = 0.0;
  for (float v : lp) sum += v;
  *loss = float(sum);
  return TG_OK;
}

int tg_mlp_adam(int device, uint32_t D, uint32_t t, float lr, const float* grad, float* params,
                float* m, float* v) {
  if (!grad || !params || !m || !v) return fail(TG_ERR_ARG, "null argument");
  if (D == 0 || D % 32 != 0 || D > 1024) return fail(TG_ERR_ARG, "D must be a multiple of 32 <= 1024");
  if (t < 1) return fail(TG_ERR_ARG, "t must be >= 1");
  HIPX(hipSetDevice(device));
  const size_t n1 = mlp_n1(D), n2 = mlp_n2(), n = n1 + n2;
  MlpScratch w;
  HIPX(hipStreamCreate(&w.st));
  HIPX(learner_reserve(w.L, D, 32, 32, 1, 1));
  fill_u32(reinterpret_cast<uint32_t*>(w.L.gpart), 0, n * w.L.splits, w.st);
  HIPX(hipMemcpyAsync(w.L.prm, params, n * 4, hipMemcpyHostToDevice, w.st));
  HIPX(hipMemcpyAsync(w.L.mom, m, n * 4, hipMemcpyHostToDevice, w.st));
  HIPX(hipMemcpyAsync(w.L.vel, v, n * 4, hipMemcpyHostToDevice, w.st));
  HIPX(hipMemcpyAsync(w.L.gpart, grad, n1 * 4, hipMemcpyHostToDevice, w.st));
  HIPX(hipMemcpyAsync(w.L.gpart + n1 * w.L.splits, grad + n1, n2 * 4, hipMemcpyHostToDevice,
                      w.st));
  LearnConfig c;
  c.lr = lr;
  launch_mlp_adam(w.L, 2, t, c, nullptr, w.st);
  if (int rc = check_launch("Adam check")) return rc;
  HIPX(hipMemcpyAsync(params, w.L.prm, n * 4, hipMemcpyDeviceToHost, w.st));
  HIPX(hipMemcpyAsync(m, w.L.mom, n * 4, hipMemcpyDeviceToHost, w.st));
  HIPX(hipMemcpyAsync(v, w.L.vel, n * 4, hipMemcpyDeviceToHost, w.st));
  HIPX(hipStreamSynchronize(w.st));
  return TG_OK;
}

int tg_get_stats(tg_state* s, tg_stats* out) {
  if (!s || !out) return fail(TG_ERR_ARG, "null argument");
  *out = s->stats;
  return TG_OK;
}

// ----------------------------------------------------------------------- kNN build
namespace {

// Learned initial ids in s->G: a sample's GNN-Descent graph trains the encoder, whose codes
// bucket every vector (learn.hip).  Adds the sample build's distance evaluations to *evals.
int learned_init(tg_state* s, const tg_knn_params* p, uint64_t* evals) {
  using clk = std::chrono::steady_clock;
  const uint32_t N = s->N, D = s->D, K = s->K;
  const uint32_t S = std::min<uint32_t>(p->sample ? p->sample : 50000u, N);
  const uint32_t Ks = std::min<uint32_t>(16u, S - 1);
  hipStream_t st = s->stream;
  const auto t0 = clk::now();
  if (s->samp_cap < S) {
    dfree(s->samp);
    if (dalloc(&s->samp, S)) return TG_ERR_HIP;
    s->samp_cap = S;
  }
  if (s->sub && s->sub->N != S) {
    tg_free(s->sub);
    s->sub = nullptr;
  }
  if (!s->sub)
    if (int rc = tg_create(S, D, s->device, &s->sub)) return rc;
  launch_sample_ids(s->samp, S, N, p->seed, st);
  launch_gather_rows(s->X, D, s->samp, S, s->sub->X, st);
  if (int rc = check_launch("learned init: sample")) return rc;
  HIPX(hipStreamSynchronize(st));
  s->sub->has_knn = s->sub->prep_valid = s->sub->has_prune = false;
  tg_knn_params sp{};
  sp.K = Ks;
  sp.iters = 4;
  sp.seed = p->seed ^ 0x5A3D1E5Bull;
  sp.precision = TG_PREC_EXACT;
  sp.init = TG_INIT_RANDOM;
  sp.flags = TG_FLAG_GRAPH_ITERS;  // short kernels: later builds replay the captured sequence
  if (int rc = tg_knn_build(s->sub, &sp)) return rc;
  *evals += s->sub->stats.knn_evals;
  s->stats.learn_ms[0] = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
  LearnConfig cfg;
  if (p->epochs) cfg.epochs = p->epochs;
  LearnStats ls;
  const hipError_t e = learned_seed(s->learn, s->X, N, D, K, s->sub->X, s->samp, S, s->sub->G,
                                    Ks, cfg, p->seed, s->G, st, &ls, s->profiling,
                                    (p->flags & TG_FLAG_GRAPH) != 0);
  if (e != hipSuccess)
    return fail(TG_ERR_HIP, std::string("learned init: ") + hipGetErrorString(e));
  s->stats.learn_ms[1] = ls.ms_train;
  s->stats.learn_ms[2] = ls.ms_encode;
  s->stats.learn_ms[3] = ls.ms_cluster;
  s->stats.learn_ms[4] = ls.ms_seed;
  s->stats.learn_loss[0] = ls.loss_first;
  s->stats.learn_loss[1] = ls.loss_last;
  s->stats.learn_buckets = ls.buckets;
  s->stats.learn_steps = ls.steps;
  s->has_codes = true;
  return TG_OK;
}

// the list buffers of one expansion iteration trade places (new lists become current)
void swap_lists(tg_state* s) {
  std::swap(s->G, s->G1);
  std::swap(s->GD, s->GD1);
  std::swap(s->R0, s->R1);
  std::swap(s->Rn0, s->Rn1);
}

}  // namespace

static int knn_build_impl(tg_state* s, const tg_knn_params* p) {
  if (!s || !p) return fail(TG_ERR_ARG, "null argument");
  const uint32_t K = p->K;
  if (K < 2 || K > 128) return fail(TG_ERR_ARG, "K must be in [2, 128]");
  if (K >= s->N) return fail(TG_ERR_ARG, "K must be smaller than N");
  // the reverse lists index destinations in blocks of at most 2^13 ids (2^13 blocks) and
  // edges with 32 bits
  if (s->N > (1u << 26) || uint64_t(s->N) * K >= 0xFFFFFFF0ull)
    return fail(TG_ERR_ARG, "kNN build: N must be at most 2^26 and N*K below 2^32");
  if (p->precision < TG_PREC_EXACT || p->precision > TG_PREC_FP8)
    return fail(TG_ERR_ARG, "unknown precision");
  if (p->precision == TG_PREC_BF16 && s->D % 8 != 0)
    return fail(TG_ERR_ARG, "BF16 screening needs D % 8 == 0");
  if (p->precision == TG_PREC_FP8 && s->D % 16 != 0)
    return fail(TG_ERR_ARG, "FP8 screening needs D % 16 == 0");
  if (p->init < TG_INIT_RANDOM || p->init > TG_INIT_GIVEN)
    return fail(TG_ERR_ARG, "unknown init");
  if (p->tile > 1) return fail(TG_ERR_ARG, "tile must be 0 or 1");
  if (p->init == TG_INIT_GIVEN && (!s->has_knn || s->K != K))
    return fail(TG_ERR_STATE, "init=given needs a current graph with the same K");
  if (s->nseg) {
    for (uint32_t c = 0; c < s->nseg; ++c)
      if (s->seg_lo_host[c + 1] - s->seg_lo_host[c] <= K)
        return fail(TG_ERR_ARG, "every segment must hold more than K rows");
    if (p->init == TG_INIT_LEARNED)
      return fail(TG_ERR_ARG, "learned init is not available on a segmented handle");
  }
  if (p->init == TG_INIT_LEARNED) {
    if (s->D % 32 != 0) return fail(TG_ERR_ARG, "learned init needs D % 32 == 0");
    if (std::min<uint32_t>(p->sample ? p->sample : 50000u, s->N) < 64)
      return fail(TG_ERR_ARG, "learned init needs a sample of at least 64 vectors");
  }
  HIPX(hipSetDevice(s->device));
  if (ensure_graph(s, K) || ensure_knn_scratch(s, K)) return TG_ERR_HIP;
  if (p->precision != TG_PREC_EXACT)
    if (int rc = ensure_prep(s)) return rc;
  if (p->precision == TG_PREC_FP8)
    if (int rc = ensure_prep8(s)) return rc;
  hipStream_t st = s->stream;
  if (ensure_events(s, 2)) return TG_ERR_HIP;
  // window events (separate from the per-section pool)
  EventPair win;
  HIPX(hipEventCreate(&win.e0));
  HIPX(hipEventCreate(&win.e1));
  hipEvent_t w0 = win.e0, w1 = win.e1;
  Sections sec{s};
  sec.next = 0;
  fill_u32(reinterpret_cast<uint32_t*>(s->counters), 0, 2 * CNT_SLOTS * CNT_N, st);
  clear_errors(s, {&s->pn, &s->po, &s->pc, &s->pr});
  HIPX(hipEventRecord(w0, st));
  // precision-dependent parts of the argument block (re-applied after every swap)
  // integer data makes the FP32 / BF16 screens exact; FP8 rounds integers above 16
  const bool exact_screen =
      (p->precision == TG_PREC_FP32 || p->precision == TG_PREC_BF16) && s->cert_int;
  const uint16_t* rows16 =
      ((p->precision == TG_PREC_BF16 || p->precision == TG_PREC_FP8) && s->cert_bf16) ? s->Xb
                                                                                      : nullptr;
  bool use_order = false;  // after a learned init: process rows in bucket order
  s->order_valid = false;
  auto args = [&]() {
    KnnArgs k = knn_args(s, p->seed);
    if (use_order) k.perm = s->learn.order;
    if (s->nseg) {
      k.seg_of = s->seg_of;
      k.seg_lo = s->seg_lo;
      k.seg_key = s->seg_key;
    }
    k.exact_screen = exact_screen ? 1 : 0;
    k.Xr16 = rows16;
    k.X8 = (p->precision != TG_PREC_EXACT && s->cert_u8 && s->D <= 256) ? s->X8 : nullptr;
    k.n2s = k.X8 ? s->n2s : nullptr;
    if (p->precision == TG_PREC_FP8) {
      k.Xq8 = s->Xq8;
      k.sc8 = s->sc8;
      k.nq = s->nq8;
      k.ne = s->ne8;
    }
    return k;
  };
  KnnArgs a = args();
  const uint32_t N = s->N;
  for (double& v : s->stats.learn_ms) v = 0.0;
  s->stats.learn_loss[0] = s->stats.learn_loss[1] = 0.f;
  s->stats.learn_buckets = s->stats.learn_steps = 0;
  uint64_t extra_evals = 0;
  if (sec.start(0)) return TG_ERR_HIP;
  if (p->init == TG_INIT_RANDOM) {
    launch_knn_init(a, st);
  } else {
    if (p->init == TG_INIT_LEARNED) {
      if (int rc = learned_init(s, p, &extra_evals)) return rc;
      use_order = true;
      a = args();
    }
    launch_knn_init_sort(a, st);
  }
  if (sec.stop()) return TG_ERR_HIP;
  const uint32_t it1 = p->iters / 2, it2 = p->iters / 2;
  // The iteration sequence.  `timed`: per-kernel sections (not while capturing a graph).
  // It ends with the handle's list buffers swapped once per expansion iteration.
  auto run_iters = [&](bool timed) -> int {
    Sections none{s};
    Sections& sc = timed ? sec : none;
    const int prof = s->profiling;
    if (!timed) s->profiling = 0;
    struct Restore {
      tg_state* s;
      int prof;
      ~Restore() { s->profiling = prof; }
    } restore{s, prof};
    for (uint32_t it = 0; it < it1; ++it) {
      if (sc.start(1)) return TG_ERR_HIP;
      pb_reset(a.pn, N, s->fill, st);
      pb_reset(a.po, N, s->fill, st);
      launch_knn_sample(a, st);
      if (sc.stop()) return TG_ERR_HIP;
      if (sc.start(2)) return TG_ERR_HIP;
      pb_group_overflow(a.pn, N, s->fill, s->scan_tmp, st);
      fill_u32(s->fill, 0, N, st);
      pb_group_overflow(a.po, N, s->fill, s->scan_tmp, st);
      launch_knn_rsample(a, st);
      if (sc.stop()) return TG_ERR_HIP;
      if (sc.start(3)) return TG_ERR_HIP;
      pb_reset(a.pc, N, s->fill, st);
      launch_knn_localjoin(a, p->precision, int(p->tile), st);
      if (sc.stop()) return TG_ERR_HIP;
      if (sc.start(4)) return TG_ERR_HIP;
      pb_group_overflow(a.pc, N, s->fill, s->scan_tmp, st);
      launch_knn_merge(a, st);
      if (sc.stop()) return TG_ERR_HIP;
      if (int rc = check_launch("GNN-Descent phase 1")) return rc;
    }
    if (p->iters == 0) return TG_OK;  // iters == 0 returns the initial lists
    if (sc.start(5)) return TG_ERR_HIP;
    if (launch_knn_rev_build(a, s->rv_hist, s->rv_scan, st))
      return fail(TG_ERR_ARG, "reverse lists: N and K beyond 2^26 ids or 2^32 edges");
    KnnArgs ar = a;
    ar.pr.cap = 0;  // the runs are the grouped overflow area
    launch_knn_rev_reduce(ar, st);
    if (sc.stop()) return TG_ERR_HIP;
    for (uint32_t it = 0; it < it2; ++it) {
      if (sc.start(6)) return TG_ERR_HIP;
      pb_reset(a.pr, N, s->fill, st);
      launch_knn_phase2(a, p->precision, st);
      pb_group_overflow(a.pr, N, s->fill, s->scan_tmp, st);
      launch_knn_rupdate(a, st);
      if (sc.stop()) return TG_ERR_HIP;
      swap_lists(s);
      a = args();
      if (int rc = check_launch("GNN-Descent phase 2")) return rc;
    }
    if (sc.start(7)) return TG_ERR_HIP;
    launch_knn_final(a, st);
    if (sc.stop()) return TG_ERR_HIP;
    return TG_OK;
  };
  if ((p->flags & TG_FLAG_GRAPH_ITERS) && p->iters > 0) {
    // key: every launch argument of the sequence
    std::vector<unsigned char> key;
    auto put = [&](const void* v, size_t n) {
      const unsigned char* b = static_cast<const unsigned char*>(v);
      key.insert(key.end(), b, b + n);
    };
    put(&a, sizeof(a));
    const uint64_t extra[] = {it1, it2, uint64_t(p->precision), p->tile,
                              uint64_t(reinterpret_cast<uintptr_t>(s->fill)),
                              uint64_t(reinterpret_cast<uintptr_t>(s->scan_tmp)),
                              uint64_t(reinterpret_cast<uintptr_t>(s->rv_hist)),
                              uint64_t(reinterpret_cast<uintptr_t>(s->rv_scan)), s->nseg};
    put(extra, sizeof(extra));
    put(s->seg_lo_host.data(), s->seg_lo_host.size() * 4);
    hipGraphExec_t exec = nullptr;
    for (auto& g : s->iter_graphs)
      if (g.first == key) exec = g.second;
    if (exec) {
      for (uint32_t it = 0; it < it2; ++it) swap_lists(s);  // what the sequence leaves
      a = args();
    } else {
      hipGraph_t g = nullptr;
      HIPX(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      const int rc = run_iters(false);
      const hipError_t ec = hipStreamEndCapture(st, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (ec != hipSuccess)
        return fail(TG_ERR_HIP, std::string("capture of the iterations: ") + hipGetErrorString(ec));
      const hipError_t ei = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ei != hipSuccess)
        return fail(TG_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
      if (s->iter_graphs.size() >= 2) {
        (void)hipGraphExecDestroy(s->iter_graphs.front().second);
        s->iter_graphs.erase(s->iter_graphs.begin());
      }
      s->iter_graphs.push_back({std::move(key), exec});
    }
    HIPX(hipGraphLaunch(exec, st));
  } else {
    if (int rc = run_iters(true)) return rc;
  }
  HIPX(hipEventRecord(w1, st));
  HIPX(hipStreamSynchronize(st));
  if (int rc = check_launch("GNN-Descent")) return rc;
  float ms = 0.f;
  HIPX(hipEventElapsedTime(&ms, w0, w1));
  s->stats.knn_ms = ms;
  for (double& v : s->stats.knn_kernel_ms) v = 0.0;
  if (sec.collect(s->stats.knn_kernel_ms)) return TG_ERR_HIP;
  unsigned long long cnt[CNT_N];
  if (int rc = read_counters(s, cnt)) return rc;
  s->stats.knn_evals = cnt[0] + extra_evals;
  s->stats.knn_rescored = cnt[1];
  if (check_pb_err(s, s->pn, "reverse sampling") || check_pb_err(s, s->po, "reverse sampling") ||
      check_pb_err(s, s->pc, "local join") || check_pb_err(s, s->pr, "reverse lists"))
    return TG_ERR_CAPACITY;
  s->has_knn = true;
  s->has_prune = false;
  s->order_valid = use_order;
  return TG_OK;
}
int tg_knn_build(tg_state* s, const tg_knn_params* p) {
  return guarded([&] { return knn_build_impl(s, p); });
}

int tg_set_knn(tg_state* s, uint32_t K, const uint32_t* ids, const float* dists) {
  if (!s || !ids || !dists) return fail(TG_ERR_ARG, "null argument");
  if (K < 2 || K > 128 || K >= s->N) return fail(TG_ERR_ARG, "K must be in [2, min(128, N-1)]");
  for (size_t i = 0; i < size_t(s->N) * K; ++i)
    if (ids[i] >= s->N) return fail(TG_ERR_ARG, "kNN id out of range");
  // rows sorted by (distance, id) with distances >= 0: the pruning kernels merge them as
  // sorted lists and the local join binary-searches them
  for (uint32_t u = 0; u < s->N; ++u) {
    const float* d = dists + size_t(u) * K;
    const uint32_t* g = ids + size_t(u) * K;
    for (uint32_t j = 0; j < K; ++j) {
      if (!(d[j] >= 0.0f)) return fail(TG_ERR_ARG, "kNN distances must be >= 0 (row " +
                                                       std::to_string(u) + ")");
      if (j && (d[j] < d[j - 1] || (d[j] == d[j - 1] && g[j] < g[j - 1])))
        return fail(TG_ERR_ARG, "kNN row " + std::to_string(u) +
                                    " is not sorted by (distance, id)");
    }
  }
  HIPX(hipSetDevice(s->device));
  if (ensure_graph(s, K)) return TG_ERR_HIP;
  HIPX(hipMemcpy(s->G, ids, size_t(s->N) * K * 4, hipMemcpyHostToDevice));
  HIPX(hipMemcpy(s->GD, dists, size_t(s->N) * K * 4, hipMemcpyHostToDevice));
  s->order_valid = false;
  s->has_knn = true;
  s->has_prune = false;
  return TG_OK;
}

int tg_get_knn(tg_state* s, uint32_t* ids, float* dists) {
  if (!s) return fail(TG_ERR_ARG, "null handle");
  if (!s->has_knn) return fail(TG_ERR_STATE, "no kNN graph");
  HIPX(hipSetDevice(s->device));
  if (ids) HIPX(hipMemcpy(ids, s->G, size_t(s->N) * s->K * 4, hipMemcpyDeviceToHost));
  if (dists) HIPX(hipMemcpy(dists, s->GD, size_t(s->N) * s->K * 4, hipMemcpyDeviceToHost));
  return TG_OK;
}

// -------------------------------------------------------------------------- prune
static int prune_impl(tg_state* s, const tg_prune_params* p) {
  if (!s || !p) return fail(TG_ERR_ARG, "null argument");
  if (!s->has_knn) return fail(TG_ERR_STATE, "prune needs a kNN graph (build or set one)");
  const uint32_t K = s->K, FD = p->FD;
  const bool rank = p->kind == TG_KIND_CAGRA || p->kind == TG_KIND_DPG;
  if (p->kind < 0 || p->kind > 3) return fail(TG_ERR_ARG, "unknown index kind");
  if (rank != (p->rev_rule == TG_RULE_RANK))
    return fail(TG_ERR_ARG,
                "rule/collect combination mixes the [deg,ids] and (id,rank) layouts "
                "(1-hop needs metric 'rank'; path and 2-hop need 'dist' or 'angle')");
  if (!rank && (p->fwd_rule < 0 || p->fwd_rule > 1 || p->rev_rule < 0 || p->rev_rule > 1))
    return fail(TG_ERR_ARG, "unknown rule");
  if (FD < 1 || FD > 128) return fail(TG_ERR_ARG, "FINAL_DEGREE must be in [1, 128]");
  if (rank && FD > K) return fail(TG_ERR_ARG, "FINAL_DEGREE must not exceed K for 1-hop");
  if (p->kind == TG_KIND_PATH && (p->TOPM < 1 || p->TOPM > 128))
    return fail(TG_ERR_ARG, "TOPM must be in [1, 128]");
  if (p->kind == TG_KIND_TWOHOP && size_t(K) * K < 480)
    return fail(TG_ERR_ARG, "2-hop collect needs K*K >= 480");
  HIPX(hipSetDevice(s->device));
  const uint32_t N = s->N;
  if (s->FD_alloc < FD) {
    dfree(s->fwd_deg);
    dfree(s->fwd_ids);
    dfree(s->fwd_d);
    dfree(s->out_deg);
    dfree(s->out_ids);
    if (dalloc(&s->fwd_deg, N) || dalloc(&s->fwd_ids, size_t(N) * FD) ||
        dalloc(&s->fwd_d, size_t(N) * FD) || dalloc(&s->out_deg, N) ||
        dalloc(&s->out_ids, size_t(N) * FD))
      return TG_ERR_HIP;
    s->FD_alloc = FD;
  }
  if (!s->cpart && dalloc(&s->cpart, size_t((N + 1023) / 1024) * s->D)) return TG_ERR_HIP;
  if (ensure_common_scratch(s)) return TG_ERR_HIP;
  if (alloc_pb(s, s->rev, 64, size_t(N) * FD)) return TG_ERR_HIP;
  hipStream_t st = s->stream;
  PruneArgs a{};
  a.X = s->X;
  a.N = N;
  a.D = s->D;
  a.K = K;
  a.G = s->G;
  a.GD = s->GD;
  a.FD = FD;
  a.TOPM = p->TOPM;
  a.th = p->thre;
  a.fwd_rule = p->fwd_rule;
  a.rev_rule = p->rev_rule;
  a.ep = s->ep_dev;
  if (s->order_valid && !s->nseg) a.perm = s->learn.order;
  if (s->nseg) {
    a.ep = s->seg_ep;
    a.seg_of = s->seg_of;
    a.seg_lo = s->seg_lo;
  }
  a.fwd_deg = s->fwd_deg;
  a.fwd_ids = s->fwd_ids;
  a.fwd_d = s->fwd_d;
  a.out_deg = s->out_deg;
  a.out_ids = s->out_ids;
  a.rev = s->rev.d;
  a.counters = s->counters;
  a.phase_clock = s->profiling >= 2 ? 1 : 0;
  a.rows = s->X;
  a.rows_bf16 = 0;
  if (int rc = ensure_prep(s)) return rc;
  if (p->precision == TG_PREC_BF16 && s->cert_bf16) {
    a.rows = s->Xb;
    a.rows_bf16 = 1;
  }
  a.rows8 = s->cert_u8 ? s->X8 : nullptr;
  // DPG's pair distances on an MFMA Gram (K <= 64) with exact re-scores near the threshold
  const bool dpg_screen = p->kind == TG_KIND_DPG && K <= 64 &&
                          ((p->precision == TG_PREC_BF16 && s->D % 8 == 0) ||
                           (p->precision == TG_PREC_FP8 && s->D % 16 == 0));
  if (dpg_screen && p->precision == TG_PREC_FP8)
    if (int rc = ensure_prep8(s)) return rc;
  EventPair win;
  HIPX(hipEventCreate(&win.e0));
  HIPX(hipEventCreate(&win.e1));
  hipEvent_t w0 = win.e0, w1 = win.e1;
  Sections sec{s};
  fill_u32(reinterpret_cast<uint32_t*>(s->counters), 0, 2 * CNT_SLOTS * CNT_N, st);
  fill_u32(s->ep_dev, 0, 1, st);
  clear_errors(s, {&s->rev});
  HIPX(hipEventRecord(w0, st));
  pb_reset(a.rev, N, s->fill, st);
  if (p->kind == TG_KIND_PATH) {
    if (sec.start(0)) return TG_ERR_HIP;
    if (s->nseg) {
      for (uint32_t c = 0; c < s->nseg; ++c) {
        const uint32_t lo = s->seg_lo_host[c], n = s->seg_lo_host[c + 1] - lo;
        launch_centroid(s->X + size_t(lo) * s->D, n, s->D, s->cpart,
                        s->seg_cent + size_t(c) * s->D, st);
      }
      launch_prune(a, PH_EP, s->seg_cent, s->seg_ep, st, s->nseg);
    } else {
      launch_centroid(s->X, N, s->D, s->cpart, s->centroid, st);
      launch_prune(a, PH_EP, s->centroid, s->ep_dev, st);
    }
    if (sec.stop()) return TG_ERR_HIP;
  }
  if (sec.start(1)) return TG_ERR_HIP;
  switch (p->kind) {
    case TG_KIND_PATH: launch_prune(a, PH_FWD_PATH, nullptr, nullptr, st); break;
    case TG_KIND_TWOHOP: launch_prune(a, PH_FWD_2HOP, nullptr, nullptr, st); break;
    case TG_KIND_CAGRA: launch_prune(a, PH_CAGRA, nullptr, nullptr, st); break;
    default:
      if (dpg_screen) {
        DpgScreen q{};
        q.nrm2 = s->nrm2;
        q.nq = p->precision == TG_PREC_FP8 ? s->nq8 : s->nq;
        q.ne = p->precision == TG_PREC_FP8 ? s->ne8 : s->ne;
        q.sc8 = s->sc8;
        q.Xb = s->Xb;
        q.Xq8 = s->Xq8;
        q.eta = float((8.0 * s->D + 128.0) * 5.9604644775390625e-8);
        launch_dpg_screen(a, q, p->precision == TG_PREC_FP8 ? KP_FP8 : KP_BF16, st);
      } else {
        launch_prune(a, PH_DPG, nullptr, nullptr, st);
      }
      break;
  }
  if (sec.stop()) return TG_ERR_HIP;
  if (sec.start(3)) return TG_ERR_HIP;
  pb_group_overflow(a.rev, N, s->fill, s->scan_tmp, st);
  if (sec.stop()) return TG_ERR_HIP;
  if (sec.start(2)) return TG_ERR_HIP;
  launch_prune(a, rank ? PH_RANK_REV : PH_REV, nullptr, nullptr, st);
  if (sec.stop()) return TG_ERR_HIP;
  HIPX(hipEventRecord(w1, st));
  HIPX(hipStreamSynchronize(st));
  if (int rc = check_launch("prune")) return rc;
  float ms = 0.f;
  HIPX(hipEventElapsedTime(&ms, w0, w1));
  s->stats.prune_ms = ms;
  for (double& v : s->stats.prune_kernel_ms) v = 0.0;
  if (sec.collect(s->stats.prune_kernel_ms)) return TG_ERR_HIP;
  unsigned long long cnt[CNT_N];
  if (int rc = read_counters(s, cnt)) return rc;
  for (int i = 0; i < 15; ++i) s->stats.prune_counters[i] = cnt[i];
  if (s->nseg) {
    s->seg_ep_host.resize(s->nseg);
    if (p->kind == TG_KIND_PATH)
      HIPX(hipMemcpy(s->seg_ep_host.data(), s->seg_ep, size_t(s->nseg) * 4, hipMemcpyDeviceToHost));
    else
      std::fill(s->seg_ep_host.begin(), s->seg_ep_host.end(), 0u);
    s->ep = s->seg_ep_host[0];
  } else {
    HIPX(hipMemcpy(&s->ep, s->ep_dev, 4, hipMemcpyDeviceToHost));
  }
  if (check_pb_err(s, s->rev, "reverse edges")) return TG_ERR_CAPACITY;
  s->has_prune = true;
  s->last_prune = *p;
  return TG_OK;
}
int tg_prune(tg_state* s, const tg_prune_params* p) {
  return guarded([&] { return prune_impl(s, p); });
}

int tg_get_pruned(tg_state* s, uint32_t* deg, uint32_t* ids, uint32_t* ep) {
  if (!s) return fail(TG_ERR_ARG, "null handle");
  if (!s->has_prune) return fail(TG_ERR_STATE, "no pruned graph");
  HIPX(hipSetDevice(s->device));
  const uint32_t FD = s->last_prune.FD;
  if (deg) HIPX(hipMemcpy(deg, s->out_deg, size_t(s->N) * 4, hipMemcpyDeviceToHost));
  if (ids) HIPX(hipMemcpy(ids, s->out_ids, size_t(s->N) * FD * 4, hipMemcpyDeviceToHost));
  if (ep) *ep = s->ep;
  return TG_OK;
}

int tg_pruned_width(tg_state* s, uint32_t* FD) {
  if (!s || !FD) return fail(TG_ERR_ARG, "null argument");
  if (!s->has_prune) return fail(TG_ERR_STATE, "no pruned graph");
  *FD = s->last_prune.FD;
  return TG_OK;
}

int tg_get_forward(tg_state* s, uint32_t* deg, uint32_t* ids, float* dists) {
  if (!s) return fail(TG_ERR_ARG, "null handle");
  if (!s->has_prune) return fail(TG_ERR_STATE, "no pruned graph");
  HIPX(hipSetDevice(s->device));
  const uint32_t FD = s->last_prune.FD;
  if (deg) HIPX(hipMemcpy(deg, s->fwd_deg, size_t(s->N) * 4, hipMemcpyDeviceToHost));
  if (ids) HIPX(hipMemcpy(ids, s->fwd_ids, size_t(s->N) * FD * 4, hipMemcpyDeviceToHost));
  if (dists) HIPX(hipMemcpy(dists, s->fwd_d, size_t(s->N) * FD * 4, hipMemcpyDeviceToHost));
  return TG_OK;
}

// ------------------------------------------------------------------------ writers
static int set_segments_impl(tg_state* s, uint32_t nseg, const uint32_t* starts,
                             const uint32_t* keys) {
  if (!s) return fail(TG_ERR_ARG, "null handle");
  HIPX(hipSetDevice(s->device));
  if (nseg == 0) {
    s->nseg = 0;
    s->seg_lo_host.clear();
  } else {
    if (!starts) return fail(TG_ERR_ARG, "null starts");
    if (starts[0] != 0 || starts[nseg] != s->N)
      return fail(TG_ERR_ARG, "segment starts must run from 0 to N");
    for (uint32_t c = 0; c < nseg; ++c)
      if (starts[c + 1] <= starts[c]) return fail(TG_ERR_ARG, "segments must be non-empty");
    if (!s->seg_of && dalloc(&s->seg_of, s->N)) return TG_ERR_HIP;
    if (s->seg_cap < nseg) {
      dfree(s->seg_lo);
      dfree(s->seg_ep);
      dfree(s->seg_key);
      dfree(s->seg_cent);
      if (dalloc(&s->seg_lo, size_t(nseg) + 1) || dalloc(&s->seg_ep, nseg) ||
          dalloc(&s->seg_key, nseg) || dalloc(&s->seg_cent, size_t(nseg) * s->D))
        return TG_ERR_HIP;
      s->seg_cap = nseg;
    }
    std::vector<uint32_t> key(nseg);
    for (uint32_t c = 0; c < nseg; ++c) key[c] = keys ? keys[c] : c;
    HIPX(hipMemcpy(s->seg_key, key.data(), size_t(nseg) * 4, hipMemcpyHostToDevice));
    std::vector<uint32_t> of(s->N);
    for (uint32_t c = 0; c < nseg; ++c)
      std::fill(of.begin() + starts[c], of.begin() + starts[c + 1], c);
    HIPX(hipMemcpy(s->seg_of, of.data(), size_t(s->N) * 4, hipMemcpyHostToDevice));
    HIPX(hipMemcpy(s->seg_lo, starts, (size_t(nseg) + 1) * 4, hipMemcpyHostToDevice));
    s->nseg = nseg;
    s->seg_lo_host.assign(starts, starts + nseg + 1);
  }
  s->has_knn = false;
  s->has_prune = false;
  s->order_valid = false;
  return TG_OK;
}
int tg_set_segments(tg_state* s, uint32_t nseg, const uint32_t* starts, const uint32_t* keys) {
  return guarded([&] { return set_segments_impl(s, nseg, starts, keys); });
}

int tg_autotune_knn(tg_state* s, uint32_t K, uint64_t seed, tg_tune_result* out) {
  if (!s || !out) return fail(TG_ERR_ARG, "null argument");
  *out = tg_tune_result{};
  const bool bf16 = s->D % 8 == 0, fp8 = s->D % 16 == 0;
  const int32_t precs[7] = {TG_PREC_EXACT, TG_PREC_FP32, TG_PREC_FP32, TG_PREC_BF16,
                            TG_PREC_BF16, TG_PREC_FP8, TG_PREC_FP8};
  const uint32_t tiles[7] = {0, 0, 1, 0, 1, 0, 1};
  const int prof = s->profiling;
  s->profiling = 0;
  tg_knn_params kp{};
  kp.K = K;
  kp.iters = 2;
  kp.seed = seed;
  kp.init = TG_INIT_RANDOM;
  // warm-up (first-touch allocations, code objects)
  kp.precision = TG_PREC_EXACT;
  if (int rc = tg_knn_build(s, &kp)) {
    s->profiling = prof;
    return rc;
  }
  double best = 1e300;
  for (int c = 0; c < 7; ++c) {
    if (precs[c] == TG_PREC_BF16 && !bf16) continue;
    if (precs[c] == TG_PREC_FP8 && !fp8) continue;
    kp.precision = precs[c];
    kp.tile = tiles[c];
    if (int rc = tg_knn_build(s, &kp)) {
      s->profiling = prof;
      return rc;
    }
    const uint32_t i = out->ncand++;
    out->cand_precision[i] = precs[c];
    out->cand_tile[i] = tiles[c];
    out->cand_ms[i] = s->stats.knn_ms;
    if (s->stats.knn_ms < best) {
      best = s->stats.knn_ms;
      out->precision = precs[c];
      out->tile = tiles[c];
    }
  }
  s->profiling = prof;
  return TG_OK;
}

int tg_get_entry_points(tg_state* s, uint32_t* eps) {
  if (!s || !eps) return fail(TG_ERR_ARG, "null argument");
  if (!s->has_prune) return fail(TG_ERR_STATE, "no pruned graph");
  if (s->nseg)
    std::copy(s->seg_ep_host.begin(), s->seg_ep_host.end(), eps);
  else
    eps[0] = s->ep;
  return TG_OK;
}

static int store_index_impl(tg_state* s, const char* path, const char* index_type,
                            uint64_t nssg_seed) {
  if (s && s->nseg) return fail(TG_ERR_STATE, "index files hold one graph: clear the segments");
  if (!s || !path || !index_type) return fail(TG_ERR_ARG, "null argument");
  if (!s->has_prune) return fail(TG_ERR_STATE, "no pruned graph to store");
  const std::string t(index_type);
  int kind = -1;
  if (t == "NSG") kind = 0;
  else if (t == "Vamana") kind = 1;
  else if (t == "NSSG") kind = 2;
  else if (t == "DPG") kind = 3;
  else if (t == "CAGRA") kind = 4;
  else return fail(TG_ERR_ARG, "unknown index type " + t);
  const uint32_t N = s->N, FD = s->last_prune.FD;
  std::vector<uint32_t> deg(N), ids(size_t(N) * FD);
  if (int rc = tg_get_pruned(s, deg.data(), ids.data(), nullptr)) return rc;
  FILE* f = fopen(path, "wb");
  if (!f) return fail(TG_ERR_IO, std::string("cannot open ") + path + " for writing");
  bool ok = true;
  auto put = [&](const void* p, size_t n) { ok = ok && fwrite(p, 1, n, f) == n; };
  // the rows are packed into one buffer and written with one call (not two per row)
  std::vector<uint32_t> body;
  auto put_rows = [&](int lead, bool full_width) {
    size_t words = 0;
    for (uint32_t u = 0; u < N; ++u) words += size_t(lead) + (full_width ? FD : deg[u]);
    body.resize(words);
    uint32_t* o = body.data();
    for (uint32_t u = 0; u < N; ++u) {
      const uint32_t n = full_width ? FD : deg[u];
      for (int l = 0; l < lead; ++l) *o++ = full_width ? FD : deg[u];
      memcpy(o, &ids[size_t(u) * FD], 4ull * n);
      o += n;
    }
    put(body.data(), 4 * words);
  };
  const uint32_t width = FD, ep = s->ep;
  if (kind == 0 || kind == 2) {
    put(&width, 4);
    if (kind == 0) {
      put(&ep, 4);
    } else {
      // first 10 of a SplitMix64-driven Fisher-Yates shuffle of 0..N-1
      std::vector<uint32_t> perm(N);
      for (uint32_t i = 0; i < N; ++i) perm[i] = i;
      uint64_t x = nssg_seed;
      const uint32_t m = std::min<uint32_t>(10, N);
      for (uint32_t i = 0; i < m; ++i) {
        const uint64_t j = i + splitmix64(x) % uint64_t(N - i);
        std::swap(perm[i], perm[j]);
      }
      uint32_t eps[10] = {0};
      for (uint32_t i = 0; i < m; ++i) eps[i] = perm[i];
      const uint32_t n_ep = 10;
      put(&n_ep, 4);
      put(eps, 40);
    }
    put_rows(1, false);
  } else if (kind == 1) {
    uint64_t size = 24;
    uint32_t maxd = 0;
    for (uint32_t u = 0; u < N; ++u) {
      size += 4ull * (deg[u] + 1);
      maxd = std::max(maxd, deg[u]);
    }
    const uint64_t frozen = 0;
    put(&size, 8);
    put(&maxd, 4);
    put(&ep, 4);
    put(&frozen, 8);
    put_rows(1, false);
  } else if (kind == 3) {
    const char magic[8] = {'K', 'N', 'N', 'G', 'R', 'A', 'P', 'H'};
    const uint32_t vmaj = 2, vmin = 0;
    put(magic, 8);
    put(&vmaj, 4);
    put(&vmin, 4);
    put(&N, 4);
    put_rows(2, false);
  } else {
    put_rows(1, true);
  }
  ok = (fclose(f) == 0) && ok;
  if (!ok) return fail(TG_ERR_IO, std::string("short write to ") + path);
  return TG_OK;
}
int tg_store_index(tg_state* s, const char* path, const char* index_type, uint64_t nssg_seed) {
  return guarded([&] { return store_index_impl(s, path, index_type, nssg_seed); });
}

// ------------------------------------------------------------------------ recall
// ------------------------------------------------------------------------ search
namespace {
constexpr uint32_t SEARCH_GRID = 256 * 12;  // workgroups: 12 queries in flight per CU
constexpr uint32_t SEARCH_VCAP = 8192;      // listed visited ids per query
}  // namespace

int tg_search(tg_state* s, uint32_t nq, const float* queries, const tg_search_params* p,
              uint32_t* ids, float* dists, tg_search_stats* out) {
  if (!s || !queries || !p || !ids) return fail(TG_ERR_ARG, "null argument");
  if (!s->has_prune) return fail(TG_ERR_STATE, "search needs a pruned graph");
  if (s->nseg) return fail(TG_ERR_STATE, "search runs on a handle without segments");
  if (nq < 1) return fail(TG_ERR_ARG, "nq must be >= 1");
  if (p->L < 1 || p->L > 256) return fail(TG_ERR_ARG, "L must be in [1, 256]");
  if (p->k < 1 || p->k > p->L) return fail(TG_ERR_ARG, "k must be in [1, L]");
  const uint32_t ep = p->ep == TG_INV_ID ? s->ep : p->ep;
  if (ep >= s->N) return fail(TG_ERR_ARG, "entry point out of range");
  HIPX(hipSetDevice(s->device));
  const uint32_t N = s->N, D = s->D;
  const uint32_t words = (N + 31) / 32;
  if (!s->sbm) {
    if (dalloc(&s->sbm, size_t(SEARCH_GRID) * words) ||
        dalloc(&s->svl, size_t(SEARCH_GRID) * SEARCH_VCAP) || dalloc(&s->scnt, 2))
      return TG_ERR_HIP;
    fill_u32(s->sbm, 0, size_t(SEARCH_GRID) * words, s->stream);
  }
  float* Q = nullptr;
  uint32_t* oi = nullptr;
  float* od = nullptr;
  if (dalloc(&Q, size_t(nq) * D) || dalloc(&oi, size_t(nq) * p->k) ||
      dalloc(&od, size_t(nq) * p->k)) {
    dfree(Q);
    dfree(oi);
    return TG_ERR_HIP;
  }
  int rc = TG_OK;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMemcpyAsync(Q, queries, size_t(nq) * D * 4, hipMemcpyHostToDevice, s->stream);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) {
    fill_u32(reinterpret_cast<uint32_t*>(s->scnt), 0, 4, s->stream);
    SearchArgs a{};
    a.X = s->X;
    a.N = N;
    a.D = D;
    a.deg = s->out_deg;
    a.ids = s->out_ids;
    a.FD = s->last_prune.FD;
    a.ep = ep;
    a.Q = Q;
    a.nq = nq;
    a.L = p->L;
    a.k = p->k;
    a.seed = p->seed;
    a.bitmap = s->sbm;
    a.words = words;
    a.vlist = s->svl;
    a.vcap = SEARCH_VCAP;
    a.out_ids = oi;
    a.out_d = od;
    a.counters = s->scnt;
    e = hipEventRecord(e0, s->stream);
    launch_search(a, std::min(nq, SEARCH_GRID), s->stream);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(e1, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  }
  unsigned long long cnt[2] = {0, 0};
  if (e == hipSuccess)
    e = hipMemcpy(ids, oi, size_t(nq) * p->k * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess && dists)
    e = hipMemcpy(dists, od, size_t(nq) * p->k * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(cnt, s->scnt, sizeof(cnt), hipMemcpyDeviceToHost);
  if (e == hipSuccess && out) {
    float ms = 0.f;
    e = hipEventElapsedTime(&ms, e0, e1);
    out->ms = ms;
    out->evals = cnt[0];
    out->expansions = cnt[1];
  }
  if (e != hipSuccess) rc = fail(TG_ERR_HIP, std::string("search: ") + hipGetErrorString(e));
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  dfree(Q);
  dfree(oi);
  dfree(od);
  return rc;
}

// ------------------------------------------------------------ allocation-free search
// tg_search_noalloc runs the same kernel as tg_search on caller-owned device buffers, with
// its scratch (the two evaluation counters, the per-workgroup visited-id lists and visited
// bitmaps) in one file-scope device array instead of handle buffers allocated on first use.
// Each GPU the library runs on holds its own copy of the array until the process ends.
// Host code never passes the array's name to a kernel (on the host it is only a stand-in
// for the symbol): the device address comes from hipGetSymbolAddress, once per device with
// that device current, and is kept per device.
namespace {
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;  // 64 MiB per GPU
constexpr size_t NOALLOC_HEAD = 256;  // counters (2 x u64), padded to the alignment
constexpr int NOALLOC_MAX_DEVICES = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// Device address of `scratch` on `device` (current on the calling thread), looked up once.
int scratch_on(int device, unsigned char** out) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES)
    return fail(TG_ERR_ARG, "device index beyond the scratch table");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(TG_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                  hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<unsigned char*>(ptr);
  }
  *out = g_scratch_ptr[device];
  return TG_OK;
}
}  // namespace

int tg_search_noalloc(tg_state* s, uint32_t nq, const float* queries_dev,
                      const tg_search_params* p, uint32_t* ids_dev, float* dists_dev,
                      void* stream, tg_search_stats* out) {
  if (!s || !queries_dev || !p || !ids_dev || !dists_dev) return fail(TG_ERR_ARG, "null argument");
  if (!s->has_prune) return fail(TG_ERR_STATE, "search needs a pruned graph");
  if (s->nseg) return fail(TG_ERR_STATE, "search runs on a handle without segments");
  if (nq < 1) return fail(TG_ERR_ARG, "nq must be >= 1");
  if (p->L < 1 || p->L > 256) return fail(TG_ERR_ARG, "L must be in [1, 256]");
  if (p->k < 1 || p->k > p->L) return fail(TG_ERR_ARG, "k must be in [1, L]");
  const uint32_t ep = p->ep == TG_INV_ID ? s->ep : p->ep;
  if (ep >= s->N) return fail(TG_ERR_ARG, "entry point out of range");
  const uint32_t N = s->N, words = (N + 31) / 32;
  // per workgroup: a visited-id list of SEARCH_VCAP ids and a bitmap of N bits
  const size_t per_wg = (size_t(SEARCH_VCAP) + words) * 4;
  const size_t fit = (NOALLOC_SCRATCH_BYTES - NOALLOC_HEAD) / per_wg;
  if (fit < 1)
    return fail(TG_ERR_CAPACITY, "allocation-free search: N = " + std::to_string(N) +
                                     " needs " + std::to_string(per_wg + NOALLOC_HEAD) +
                                     " bytes of scratch per workgroup, above the " +
                                     std::to_string(NOALLOC_SCRATCH_BYTES) + "-byte maximum");
  const uint32_t grid = uint32_t(std::min<size_t>({size_t(nq), size_t(SEARCH_GRID), fit}));
  HIPX(hipSetDevice(s->device));
  unsigned char* base = nullptr;
  if (int rc = scratch_on(s->device, &base)) return rc;
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : s->stream;
  auto* counters = reinterpret_cast<unsigned long long*>(base);
  auto* vlist = reinterpret_cast<uint32_t*>(base + NOALLOC_HEAD);
  uint32_t* bitmap = vlist + size_t(grid) * SEARCH_VCAP;
  // the kernel leaves its bitmaps zero, but an interrupted earlier call may not have:
  // clear the counters and the bitmaps this launch uses (fill kernels, no memset nodes)
  fill_u32(reinterpret_cast<uint32_t*>(counters), 0, 4, st);
  fill_u32(bitmap, 0, size_t(grid) * words, st);
  SearchArgs a{};
  a.X = s->X;
  a.N = N;
  a.D = s->D;
  a.deg = s->out_deg;
  a.ids = s->out_ids;
  a.FD = s->last_prune.FD;
  a.ep = ep;
  a.Q = queries_dev;
  a.nq = nq;
  a.L = p->L;
  a.k = p->k;
  a.seed = p->seed;
  a.bitmap = bitmap;
  a.words = words;
  a.vlist = vlist;
  a.vcap = SEARCH_VCAP;
  a.out_ids = ids_dev;
  a.out_d = dists_dev;
  a.counters = counters;
  launch_search(a, grid, st);
  HIPX(hipGetLastError());
  if (out) {
    unsigned long long cnt[2] = {0, 0};
    HIPX(hipStreamSynchronize(st));
    HIPX(hipMemcpy(cnt, counters, sizeof(cnt), hipMemcpyDeviceToHost));
    out->ms = 0.0;
    out->evals = cnt[0];
    out->expansions = cnt[1];
  }
  return TG_OK;
}

int tg_search_recall(tg_state* s, uint32_t nq, const float* queries, uint32_t k,
                     const uint32_t* ids, double* recall) {
  if (!s || !queries || !ids || !recall) return fail(TG_ERR_ARG, "null argument");
  if (k < 1 || k > 16) return fail(TG_ERR_ARG, "k must be in [1, 16]");
  if (nq < 1) return fail(TG_ERR_ARG, "nq must be >= 1");
  HIPX(hipSetDevice(s->device));
  const uint32_t cap = 256;
  float* Q = nullptr;
  uint32_t *G = nullptr, *cand_n = nullptr, *hits = nullptr;
  double *kth = nullptr, *cand_d = nullptr;
  int rc = TG_OK;
  if (dalloc(&Q, size_t(nq) * s->D) || dalloc(&G, size_t(nq) * k) || dalloc(&cand_n, nq) ||
      dalloc(&hits, nq) || dalloc(&kth, nq) ||
      dalloc(&cand_d, size_t(nq) * cap + size_t(nq) * k)) {
    rc = TG_ERR_HIP;
  } else {
    hipError_t e = hipMemcpy(Q, queries, size_t(nq) * s->D * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(G, ids, size_t(nq) * k * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      launch_recall(s->X, s->N, s->D, G, k, nullptr, nq, k, kth, cand_n, cand_d, cap, hits,
                    s->stream, Q);
      e = hipStreamSynchronize(s->stream);
    }
    if (e == hipSuccess) e = hipGetLastError();
    std::vector<uint32_t> h(nq);
    if (e == hipSuccess) e = hipMemcpy(h.data(), hits, size_t(nq) * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      rc = fail(TG_ERR_HIP, std::string("search recall: ") + hipGetErrorString(e));
    } else {
      double tot = 0.0;
      for (uint32_t i = 0; i < nq; ++i) tot += double(std::min(h[i], k)) / k;
      *recall = tot / nq;
    }
  }
  dfree(Q);
  dfree(G);
  dfree(cand_n);
  dfree(hits);
  dfree(kth);
  dfree(cand_d);
  return rc;
}

// judged nodes: the first nq entries of a SplitMix64 Fisher-Yates shuffle of 0..N-1
static std::vector<uint32_t> judged_nodes(uint32_t N, uint32_t nq, uint64_t seed) {
  std::vector<uint32_t> perm(N);
  for (uint32_t i = 0; i < N; ++i) perm[i] = i;
  uint64_t x = seed;
  for (uint32_t i = 0; i < nq; ++i) {
    const uint64_t j = i + splitmix64(x) % uint64_t(N - i);
    std::swap(perm[i], perm[j]);
  }
  perm.resize(nq);
  return perm;
}

static int recall_impl(tg_state* s, uint32_t nq, uint64_t seed, uint32_t k, double* recall) {
  if (s && s->nseg) return fail(TG_ERR_STATE, "recall is judged on a handle without segments");
  if (!s || !recall) return fail(TG_ERR_ARG, "null argument");
  if (!s->has_knn) return fail(TG_ERR_STATE, "no kNN graph");
  if (k < 1 || k > 16 || k > s->K) return fail(TG_ERR_ARG, "k must be in [1, min(16, K)]");
  if (nq < 1 || nq > s->N) return fail(TG_ERR_ARG, "nq must be in [1, N]");
  HIPX(hipSetDevice(s->device));
  const std::vector<uint32_t> perm = judged_nodes(s->N, nq, seed);
  const uint32_t cap = 256;
  uint32_t *q = nullptr, *cand_n = nullptr, *hits = nullptr;
  double *kth = nullptr, *cand_d = nullptr;
  int rc = TG_OK;
  if (dalloc(&q, nq) || dalloc(&cand_n, nq) || dalloc(&hits, nq) || dalloc(&kth, nq) ||
      dalloc(&cand_d, size_t(nq) * cap + size_t(nq) * k)) {
    rc = TG_ERR_HIP;
  } else {
    hipError_t e = hipMemcpy(q, perm.data(), size_t(nq) * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      launch_recall(s->X, s->N, s->D, s->G, s->K, q, nq, k, kth, cand_n, cand_d, cap, hits,
                    s->stream);
      e = hipStreamSynchronize(s->stream);
    }
    if (e == hipSuccess) e = hipGetLastError();
    std::vector<uint32_t> h(nq);
    if (e == hipSuccess) e = hipMemcpy(h.data(), hits, size_t(nq) * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      rc = fail(TG_ERR_HIP, std::string("recall: ") + hipGetErrorString(e));
    } else {
      double tot = 0.0;
      for (uint32_t i = 0; i < nq; ++i) tot += double(std::min(h[i], k)) / k;
      *recall = tot / nq;
    }
  }
  dfree(q);
  dfree(cand_n);
  dfree(hits);
  dfree(kth);
  dfree(cand_d);
  return rc;
}
int tg_recall(tg_state* s, uint32_t nq, uint64_t seed, uint32_t k, double* recall) {
  return guarded([&] { return recall_impl(s, nq, seed, k, recall); });
}

static int dist_check_impl(tg_state* s, uint32_t nq, uint64_t seed, double* max_rel_err) {
  if (s && s->nseg) return fail(TG_ERR_STATE, "distances are judged on a handle without segments");
  if (!s || !max_rel_err) return fail(TG_ERR_ARG, "null argument");
  if (!s->has_knn) return fail(TG_ERR_STATE, "no kNN graph");
  if (nq < 1 || nq > s->N) return fail(TG_ERR_ARG, "nq must be in [1, N]");
  HIPX(hipSetDevice(s->device));
  const std::vector<uint32_t> nodes = judged_nodes(s->N, nq, seed);
  uint32_t* q = nullptr;
  unsigned long long* mx = nullptr;
  int rc = TG_OK;
  if (dalloc(&q, nq) || dalloc(&mx, 1)) {
    rc = TG_ERR_HIP;
  } else {
    hipError_t e = hipMemcpy(q, nodes.data(), size_t(nq) * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      launch_dist_check(s->X, s->N, s->D, s->G, s->GD, s->K, q, nq, mx, s->stream);
      e = hipStreamSynchronize(s->stream);
    }
    if (e == hipSuccess) e = hipGetLastError();
    unsigned long long bits = 0;
    if (e == hipSuccess) e = hipMemcpy(&bits, mx, 8, hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = fail(TG_ERR_HIP, std::string("distance check: ") + hipGetErrorString(e));
    else memcpy(max_rel_err, &bits, 8);
  }
  dfree(q);
  dfree(mx);
  return rc;
}
int tg_knn_dist_check(tg_state* s, uint32_t nq, uint64_t seed, double* max_rel_err) {
  return guarded([&] { return dist_check_impl(s, nq, seed, max_rel_err); });
}

int tg_buffer(tg_state* s, int32_t idx, void** dev_ptr, size_t* bytes) {
  if (!s || !dev_ptr || !bytes) return fail(TG_ERR_ARG, "null argument");
  const size_t N = s->N, K = s->K;
  switch (idx) {
    case 0: *dev_ptr = s->G; *bytes = N * K * 4; break;
    case 1: *dev_ptr = s->R0; *bytes = s->R0 ? N * K * 8 : 0; break;
    case 2: *dev_ptr = s->X; *bytes = N * s->D * 4; break;
    case 3: *dev_ptr = s->GD; *bytes = N * K * 4; break;
    case 4: *dev_ptr = s->R1; *bytes = s->R1 ? N * K * 8 : 0; break;
    case 5: *dev_ptr = s->Rn0; *bytes = s->Rn0 ? N * 4 : 0; break;
    case 6: *dev_ptr = s->centroid; *bytes = size_t(s->D) * 4; break;
    case 7: *dev_ptr = s->H; *bytes = s->H ? N * 128 * 4 : 0; break;
    default: return fail(TG_ERR_ARG, "buffer index must be in [0, 7]");
  }
  return TG_OK;
}

}  // extern "C"
