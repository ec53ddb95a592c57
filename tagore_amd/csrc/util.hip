// util.hip -- fills, scans, push-buffer bookkeeping, synthetic data and the FP64 recall
// judge.  None of these is on the measured hot path except the push-buffer bookkeeping.
#include "common.hpp"
#include "kernels.hpp"

namespace tg {

// ----------------------------------------------------------------------------- fills
__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_u8(uint8_t* p, uint8_t v, size_t n) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x)
    p[i] = v;
}
static unsigned grid_for(size_t n) {
  size_t g = (n + 255) / 256;
  if (g > 8192) g = 8192;
  if (g == 0) g = 1;
  return unsigned(g);
}
void fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s) {
  k_fill_u32<<<grid_for(n), 256, 0, s>>>(p, v, n);
}
void fill_u8(uint8_t* p, uint8_t v, size_t n, hipStream_t s) {
  k_fill_u8<<<grid_for(n), 256, 0, s>>>(p, v, n);
}

// ------------------------------------------------------------------------------ scans
// Exclusive scan of c[v] = max(0, cnt[v] - cap) into off[0..N]; off[N] = total.
constexpr int SCAN_BLOCK = 1024;  // elements per block (256 threads x 4)

__device__ int block_excl_scan_256(int v, int* sh) {
  // sh: 256 ints
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int s = 1; s < 256; s <<= 1) {
    int o = tid >= s ? sh[tid - s] : 0;
    __syncthreads();
    sh[tid] += o;
    __syncthreads();
  }
  const int incl = sh[tid];
  __syncthreads();
  return incl - v;
}

__global__ void k_scan1(const uint32_t* cnt, uint32_t cap, uint32_t N, uint32_t* off,
                        uint32_t* bsum) {
  __shared__ int sh[256];
  const size_t base = size_t(blockIdx.x) * SCAN_BLOCK + threadIdx.x * 4;
  int v[4], tot = 0;
  for (int q = 0; q < 4; ++q) {
    const size_t i = base + q;
    const uint32_t c = i < N ? cnt[i] : 0;
    v[q] = c > cap ? int(c - cap) : 0;
    tot += v[q];
  }
  int ex = block_excl_scan_256(tot, sh);
  for (int q = 0; q < 4; ++q) {
    const size_t i = base + q;
    if (i < N) off[i] = ex;
    ex += v[q];
  }
  if (threadIdx.x == 255) bsum[blockIdx.x] = ex;
}

__global__ void k_scan2(uint32_t* bsum, uint32_t nb, uint32_t* off, uint32_t N) {
  __shared__ int sh[256];
  int carry = 0;
  for (uint32_t base = 0; base < nb; base += 256) {
    const uint32_t i = base + threadIdx.x;
    const int v = i < nb ? int(bsum[i]) : 0;
    const int ex = block_excl_scan_256(v, sh);
    if (i < nb) bsum[i] = carry + ex;
    // total of this chunk
    __shared__ int tot;
    if (threadIdx.x == 255) tot = ex + v;
    __syncthreads();
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) off[N] = carry;
}

__global__ void k_scan3(uint32_t* off, const uint32_t* bsum, uint32_t N) {
  const size_t base = size_t(blockIdx.x) * SCAN_BLOCK + threadIdx.x * 4;
  const uint32_t add = bsum[blockIdx.x];
  for (int q = 0; q < 4; ++q) {
    const size_t i = base + q;
    if (i < N) off[i] += add;
  }
}

size_t scan_tmp_words(uint32_t N) { return (size_t(N) + SCAN_BLOCK - 1) / SCAN_BLOCK + 1; }

void excl_scan_u32(uint32_t* v, uint32_t n, uint32_t* tmp, hipStream_t s) {
  const uint32_t nb = (n + SCAN_BLOCK - 1) / SCAN_BLOCK;
  k_scan1<<<nb, 256, 0, s>>>(v, 0, n, v, tmp);  // each thread reads its elements before writing
  k_scan2<<<1, 256, 0, s>>>(tmp, nb, v, n);
  k_scan3<<<nb, 256, 0, s>>>(v, tmp, n);
}

__global__ void k_ov_scatter(PushBuf pb, uint32_t* fill) {
  const uint32_t used = min(*pb.ov_used, pb.ov_cap);
  for (uint32_t o = blockIdx.x * blockDim.x + threadIdx.x; o < used; o += gridDim.x * blockDim.x) {
    const uint32_t v = pb.ov_dest[o];
    const uint32_t dst = pb.ov_off[v] + atomicAdd(&fill[v], 1u);
    // the grouped offsets count every overflowing push; when the overflow area itself was
    // exhausted (*err set) some of them were dropped and dst may exceed the area
    if (dst < pb.ov_cap) pb.ov_sorted[dst] = pb.ov_key[o];
  }
}

__global__ void k_pb_reset_small(PushBuf pb) {
  if (threadIdx.x == 0) {
    *pb.ov_used = 0;
  }
}

void pb_reset(const PushBuf& pb, uint32_t N, uint32_t* fill, hipStream_t s) {
  fill_u32(pb.cnt, 0, N, s);
  fill_u32(fill, 0, N, s);
  k_pb_reset_small<<<1, 64, 0, s>>>(pb);
}

void pb_group_overflow(const PushBuf& pb, uint32_t N, uint32_t* fill, uint32_t* scan_tmp,
                       hipStream_t s) {
  const uint32_t nb = (N + SCAN_BLOCK - 1) / SCAN_BLOCK;
  k_scan1<<<nb, 256, 0, s>>>(pb.cnt, pb.cap, N, pb.ov_off, scan_tmp);
  k_scan2<<<1, 256, 0, s>>>(scan_tmp, nb, pb.ov_off, N);
  k_scan3<<<nb, 256, 0, s>>>(pb.ov_off, scan_tmp, N);
  // a grid-stride loop of returning atomics: 8 workgroups per CU keep enough waves in flight
  k_ov_scatter<<<2048, 256, 0, s>>>(pb, fill);
}

// ---------------------------------------------------------------------- synthetic data
// Counter-based generator: every value is a pure function of (seed, stream, index), built on
// the SplitMix64 finaliser (Steele, Lea, Flood 2014) and Box-Muller normals.
__device__ __forceinline__ uint64_t mixctr(uint64_t key, uint64_t ctr) {
  uint64_t z = key + (ctr + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double u01(uint64_t x) { return double(x >> 11) * 0x1.0p-53; }
__device__ __forceinline__ double gnormal(uint64_t key, uint64_t ctr) {
  const double u1 = 1.0 - u01(mixctr(key, 2 * ctr));
  const double u2 = u01(mixctr(key, 2 * ctr + 1));
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}
constexpr uint64_t KEY_CENTRE = 0xC3A5C85C97CB3127ull;
constexpr uint64_t KEY_ASSIGN = 0xB492B66FBE98F273ull;
constexpr uint64_t KEY_NOISE = 0x9AE16A3B2F90404Full;

__global__ void k_gen_centres(int kind, uint32_t ncomp, uint32_t D, uint64_t seed, float* C) {
  const size_t n = size_t(ncomp) * D;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x) {
    if (kind == 0)
      C[i] = float(255.0 * u01(mixctr(seed ^ KEY_CENTRE, i)));
    else
      C[i] = float(gnormal(seed ^ KEY_CENTRE, i));
  }
}

__global__ void k_gen_points(int kind, uint32_t N, uint32_t D, uint32_t ncomp, uint64_t seed,
                             float sigma, const float* C, float* X, uint32_t* maxbits) {
  const size_t n = size_t(N) * D;
  uint32_t mx = 0;
  for (size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
       e += size_t(gridDim.x) * blockDim.x) {
    const size_t i = e / D, j = e % D;
    const uint32_t c = uint32_t((mixctr(seed ^ KEY_ASSIGN, i) >> 32) * uint64_t(ncomp) >> 32);
    const double z = gnormal(seed ^ KEY_NOISE, e);
    double x = double(C[size_t(c) * D + j]);
    if (kind == 0) {
      x = rint(fmin(fmax(x + double(sigma) * z, 0.0), 255.0));
    } else if (kind == 1) {
      x = x + z;
    } else {
      x = fabs(x + z);
    }
    const float xf = float(x);
    X[e] = xf;
    if (kind == 2) mx = max(mx, __float_as_uint(xf));
  }
  if (kind == 2) atomicMax(maxbits, mx);
}

// unit L2 rows: one wave per row, FP64 sum of squares
__global__ void k_normalize_rows(float* X, uint32_t N, uint32_t D) {
  const uint32_t row = blockIdx.x;
  if (row >= N) return;
  float* x = X + size_t(row) * D;
  double s = 0.0;
  for (uint32_t j = threadIdx.x; j < D; j += TG_WAVE) s += double(x[j]) * double(x[j]);
  for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, TG_WAVE);
  const double nrm = sqrt(s);
  for (uint32_t j = threadIdx.x; j < D; j += TG_WAVE) x[j] = float(double(x[j]) / nrm);
}

__global__ void k_scale_by_max(float* X, size_t n, const uint32_t* maxbits) {
  const float m = __uint_as_float(*maxbits);
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x)
    X[i] = X[i] / m;
}

void launch_gen_data(int kind, uint32_t N, uint32_t D, uint32_t ncomp, uint64_t seed,
                     float sigma, float* centres, float* X, uint32_t* scratch, hipStream_t s) {
  k_gen_centres<<<grid_for(size_t(ncomp) * D), 256, 0, s>>>(kind, ncomp, D, seed, centres);
  fill_u32(scratch, 0, 1, s);
  k_gen_points<<<grid_for(size_t(N) * D), 256, 0, s>>>(kind, N, D, ncomp, seed, sigma, centres,
                                                       X, scratch);
  if (kind == 1) k_normalize_rows<<<N, TG_WAVE, 0, s>>>(X, N, D);
  if (kind == 2) k_scale_by_max<<<grid_for(size_t(N) * D), 256, 0, s>>>(X, size_t(N) * D, scratch);
}

// ------------------------------------------------------------------- FP64 recall judge
// d64(a,b) = sum over i ascending of (double(a_i) - double(b_i))^2, one rounding per step
// (compiled without contraction), identical to the CPU oracle's knn_f64.
__device__ double d64_rows(const float* a, const float* b, uint32_t D) {
  double s = 0.0;
  for (uint32_t i = 0; i < D; ++i) {
    const double t = double(a[i]) - double(b[i]);
    s += t * t;
  }
  return s;
}

// Queries are base rows q[i] (then the row itself is excluded and the graph row G[q[i]]
// is judged) or, with Q, external rows Q[i] (nothing excluded, G row i is judged).
__device__ __forceinline__ const float* judge_row(const float* X, const float* Q,
                                                  const uint32_t* q, uint32_t qi, uint32_t D) {
  return Q ? Q + size_t(qi) * D : X + size_t(q[qi]) * D;
}

// pass 1: the graph's first k valid entries bound the k-th exact distance from above
__global__ void k_recall_bound(const float* X, uint32_t D, const uint32_t* G, uint32_t K,
                               const uint32_t* q, uint32_t nq, uint32_t k, double* bound,
                               double* gd64, uint32_t N, const float* Q) {
  const uint32_t qi = blockIdx.x * blockDim.x + threadIdx.x;
  if (qi >= nq) return;
  const uint32_t u = Q ? 0xFFFFFFFFu : q[qi];
  const uint32_t grow = Q ? qi : u;
  const float* xq = judge_row(X, Q, q, qi, D);
  double b = 0.0;
  uint32_t nvalid = 0;
  for (uint32_t j = 0; j < k; ++j) {
    const uint32_t v = G[size_t(grow) * K + j];
    double d = __longlong_as_double(0x7FF0000000000000ll);
    if (v < N && v != u) {
      d = d64_rows(xq, X + size_t(v) * D, D);
      ++nvalid;
      b = fmax(b, d);
    }
    gd64[size_t(qi) * k + j] = d;
  }
  bound[qi] = nvalid == k ? b : __longlong_as_double(0x7FF0000000000000ll);
}

// pass 2: collect every point within the bound (tile of QT queries per block)
template <int QT>
__global__ __launch_bounds__(256) void k_recall_collect(const float* X, uint32_t N, uint32_t D,
                                                        const uint32_t* q, uint32_t nq,
                                                        const double* bound, uint32_t* cand_n,
                                                        double* cand_d, uint32_t cap,
                                                        const float* Q) {
  extern __shared__ float qrows[];  // QT * D
  const uint32_t q0 = blockIdx.y * QT;
  for (uint32_t e = threadIdx.x; e < QT * D; e += blockDim.x) {
    const uint32_t qq = e / D;
    qrows[e] = (q0 + qq < nq) ? judge_row(X, Q, q, q0 + qq, D)[e % D] : 0.f;
  }
  __syncthreads();
  double bnd[QT];
  uint32_t qid[QT];
  for (int qq = 0; qq < QT; ++qq) {
    bnd[qq] = (q0 + qq < nq) ? bound[q0 + qq] : -1.0;
    qid[qq] = (q0 + qq < nq && !Q) ? q[q0 + qq] : 0xFFFFFFFFu;
  }
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < N; p += gridDim.x * blockDim.x) {
    double s[QT];
    for (int qq = 0; qq < QT; ++qq) s[qq] = 0.0;
    const float* x = X + size_t(p) * D;
    for (uint32_t i = 0; i < D; ++i) {
      const double xi = double(x[i]);
      for (int qq = 0; qq < QT; ++qq) {
        const double t = double(qrows[qq * D + i]) - xi;
        s[qq] += t * t;
      }
    }
    for (int qq = 0; qq < QT; ++qq) {
      if (p != qid[qq] && s[qq] <= bnd[qq]) {
        const uint32_t slot = atomicAdd(&cand_n[q0 + qq], 1u);
        if (slot < cap) cand_d[size_t(q0 + qq) * cap + slot] = s[qq];
      }
    }
  }
}

// pass 3: k-th smallest candidate; hits of the graph's first k entries
__global__ void k_recall_final(uint32_t nq, uint32_t k, const uint32_t* cand_n, double* cand_d,
                               uint32_t cap, const double* gd64, double* kth, uint32_t* hits) {
  const uint32_t qi = blockIdx.x * blockDim.x + threadIdx.x;
  if (qi >= nq) return;
  const uint32_t n = cand_n[qi];
  if (n > cap || n < k) {
    kth[qi] = __longlong_as_double(0x7FF8000000000000ll);  // NaN: resolved by the fallback
    hits[qi] = 0xFFFFFFFFu;
    return;
  }
  double* c = cand_d + size_t(qi) * cap;
  // partial selection sort of the k smallest
  for (uint32_t a = 0; a < k; ++a) {
    uint32_t m = a;
    for (uint32_t b = a + 1; b < n; ++b)
      if (c[b] < c[m]) m = b;
    const double tmp = c[a];
    c[a] = c[m];
    c[m] = tmp;
  }
  const double kk = c[k - 1];
  kth[qi] = kk;
  uint32_t h = 0;
  for (uint32_t j = 0; j < k; ++j) h += gd64[size_t(qi) * k + j] <= kk ? 1u : 0u;
  hits[qi] = h;
}

// fallback for queries whose candidate list overflowed: full top-k scan, one block per query
__global__ __launch_bounds__(256) void k_recall_full(const float* X, uint32_t N, uint32_t D,
                                                     const uint32_t* q, uint32_t nq, uint32_t k,
                                                     const double* gd64, double* kth,
                                                     uint32_t* hits, const float* Q) {
  const uint32_t qi = blockIdx.x;
  if (qi >= nq || hits[qi] != 0xFFFFFFFFu) return;
  __shared__ double best[256 * 16];
  double loc[16];
  for (int j = 0; j < 16; ++j) loc[j] = __longlong_as_double(0x7FF0000000000000ll);
  const uint32_t u = Q ? 0xFFFFFFFFu : q[qi];
  const float* xq = judge_row(X, Q, q, qi, D);
  for (uint32_t p = threadIdx.x; p < N; p += blockDim.x) {
    if (p == u) continue;
    double d = d64_rows(xq, X + size_t(p) * D, D);
    if (d < loc[k - 1]) {
      int j = int(k) - 1;
      while (j > 0 && loc[j - 1] > d) {
        loc[j] = loc[j - 1];
        --j;
      }
      loc[j] = d;
    }
  }
  for (uint32_t j = 0; j < k; ++j) best[threadIdx.x * k + j] = loc[j];
  __syncthreads();
  if (threadIdx.x == 0) {
    // k-way selection over 256 sorted lists
    uint32_t ptr[256];
    for (int i = 0; i < 256; ++i) ptr[i] = 0;
    double v = 0.0;
    for (uint32_t r = 0; r < k; ++r) {
      int bi = 0;
      double bv = __longlong_as_double(0x7FF0000000000000ll);
      for (int i = 0; i < 256; ++i)
        if (ptr[i] < k && best[i * k + ptr[i]] < bv) {
          bv = best[i * k + ptr[i]];
          bi = i;
        }
      ptr[bi]++;
      v = bv;
    }
    kth[qi] = v;
    uint32_t h = 0;
    for (uint32_t j = 0; j < k; ++j) h += gd64[size_t(qi) * k + j] <= v ? 1u : 0u;
    hits[qi] = h;
  }
}

void launch_recall(const float* X, uint32_t N, uint32_t D, const uint32_t* G, uint32_t K,
                   const uint32_t* q, uint32_t nq, uint32_t k, double* kth, uint32_t* cand_n,
                   double* cand_d, uint32_t cand_cap, uint32_t* hits, hipStream_t s,
                   const float* Q) {
  // kth doubles as the bound buffer in pass 1/2; gd64 lives after cand_d
  double* bound = kth;
  double* gd64 = cand_d + size_t(nq) * cand_cap;
  k_recall_bound<<<(nq + 255) / 256, 256, 0, s>>>(X, D, G, K, q, nq, k, bound, gd64, N, Q);
  fill_u32(cand_n, 0, nq, s);
  if (D <= 256) {
    constexpr int QT = 16;
    dim3 grid(64, (nq + QT - 1) / QT);
    k_recall_collect<QT><<<grid, 256, QT * D * sizeof(float), s>>>(X, N, D, q, nq, bound, cand_n,
                                                                  cand_d, cand_cap, Q);
  } else {
    constexpr int QT = 4;
    dim3 grid(64, (nq + QT - 1) / QT);
    k_recall_collect<QT><<<grid, 256, QT * D * sizeof(float), s>>>(X, N, D, q, nq, bound, cand_n,
                                                                  cand_d, cand_cap, Q);
  }
  k_recall_final<<<(nq + 255) / 256, 256, 0, s>>>(nq, k, cand_n, cand_d, cand_cap, gd64, kth,
                                                  hits);
  k_recall_full<<<nq, 256, 0, s>>>(X, N, D, q, nq, k, gd64, kth, hits, Q);
}

// Stored FP32 kNN distances against FP64: for each sampled node u = q[i] (one block), every
// valid entry v of its row gives |GD[u][j] - d64(u, v)| / d64(u, v) (0 when both are 0, +inf
// when only the FP64 distance is 0); the largest goes to *max_bits as the bits of a
// non-negative double (their order is the values' order).
__global__ __launch_bounds__(128) void k_dist_check(const float* X, uint32_t N, uint32_t D,
                                                    const uint32_t* G, const float* GD,
                                                    uint32_t K, const uint32_t* q,
                                                    unsigned long long* max_bits) {
  const uint32_t u = q[blockIdx.x];
  for (uint32_t j = threadIdx.x; j < K; j += blockDim.x) {
    const uint32_t v = G[size_t(u) * K + j];
    if (v >= N) continue;
    const double d = d64_rows(X + size_t(u) * D, X + size_t(v) * D, D);
    const double st = double(GD[size_t(u) * K + j]);
    double rel = 0.0;
    if (d > 0.0) rel = fabs(st - d) / d;
    else if (st != 0.0) rel = __longlong_as_double(0x7FF0000000000000ll);
    atomicMax(max_bits, (unsigned long long)__double_as_longlong(rel));
  }
}

void launch_dist_check(const float* X, uint32_t N, uint32_t D, const uint32_t* G,
                       const float* GD, uint32_t K, const uint32_t* q, uint32_t nq,
                       unsigned long long* max_bits, hipStream_t s) {
  fill_u32(reinterpret_cast<uint32_t*>(max_bits), 0u, 2, s);
  k_dist_check<<<nq, 128, 0, s>>>(X, N, D, G, GD, K, q, max_bits);
}

// ------------------------------------------------------- register sort self-test
// One wave per group of 64*NPL keys.  mode 0: reg_bitonic; mode 1 (NPL 1): the 64 keys form
// a bitonic sequence, reg_bitonic_merge64 sorts it; mode 2 (NPL 2): slots 0 and 1 hold two
// ascending lists A and B (key l of each in lane l), the output's slot 0 is the 64 smallest
// keys of A and B, ascending -- the beam update of the prune kernels.
template <int NPL>
__global__ __launch_bounds__(64) void k_selftest_sort(const uint64_t* in, uint64_t* out, int mode) {
  const int lane = threadIdx.x;
  const size_t base = size_t(blockIdx.x) * 64 * NPL;
  uint64_t v[NPL];
#pragma unroll
  for (int r = 0; r < NPL; ++r) v[r] = in[base + r * 64 + lane];
  if (mode == 0) {
    reg_bitonic<NPL>(v);
  } else if (mode == 1) {
    reg_bitonic_merge64(v[0]);
  } else if (NPL >= 2) {
    const uint64_t b = lane_xor64(v[1], 63);
    v[0] = v[0] < b ? v[0] : b;
    reg_bitonic_merge64(v[0]);
  }
#pragma unroll
  for (int r = 0; r < NPL; ++r) out[base + r * 64 + lane] = v[r];
}

void launch_selftest_sort(const uint64_t* in, uint64_t* out, int npl, int mode, uint32_t groups,
                          hipStream_t s) {
  if (npl == 1) k_selftest_sort<1><<<groups, 64, 0, s>>>(in, out, mode);
  else if (npl == 2) k_selftest_sort<2><<<groups, 64, 0, s>>>(in, out, mode);
  else k_selftest_sort<8><<<groups, 64, 0, s>>>(in, out, mode);
}

}  // namespace tg
