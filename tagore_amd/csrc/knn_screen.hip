// knn_screen.hip -- screened local join on MFMA (FP32 or BF16) with exact FP32 re-scoring.
//
// The local join of GNN-Descent (nn_descent_opt_cal, Tagore_src.cu:684-1033) needs, for every
// sampled node list A = new'(u) (<= 64 rows) and B = old'(u) (<= 64 rows), the per-16-wide-
// tile minimum of the pair distances A x B, B x A and A x A.  Here the Gram matrices A.B^T and
// A.A^T are contracted on the matrix cores from LDS-staged, coalesced row loads:
//   BF16: v_mfma_f32_16x16x32_bf16 on a BF16 copy of the vectors,
//   FP32: v_mfma_f32_16x16x4_f32 on the FP32 vectors,
// and turned into screen distances s = |a|^2 + |b|^2 - 2 a'.b'.  A rigorous bound E(a,b) on
// |s - dist(a,b)| (quantisation + accumulation + canonical-distance rounding) selects, per
// tile, the only candidates that can be the exact (distance, id) minimum and pass the
// destination's worst distance; those are re-scored with the canonical FP32 distance.  The
// pushes are therefore bit-identical to the exact kernel's: graph decisions stay FP32.
#include "common.hpp"
#include "kernels.hpp"

namespace tg {


constexpr int LJ_TILE = 16;
constexpr int HROW_S = 128;


// Per-vector screen statistics (one wave per row, FP64 sums) and data-set certificates:
// flags[0] |= some value is not exactly representable in BF16,
// flags[1] |= some value is not an integer in [-256, 256], flags[2] |= some value is
// negative, flags[3] = max |x|^2 (FP32 bits; exact for integer data below 2^24),
// flags[4] |= some value is above 255.
__global__ void k_screen_prep(const float* X, uint32_t N, uint32_t D, uint16_t* Xb, float* nrm2,
                              float* nq, float* ne, uint32_t* flags) {
  const uint32_t row = blockIdx.x;
  if (row >= N) return;
  const float* x = X + size_t(row) * D;
  double s2 = 0.0, q2 = 0.0, e2 = 0.0;
  uint32_t not_bf16 = 0, not_int = 0, neg = 0, big = 0;
  for (uint32_t j = threadIdx.x; j < D; j += TG_WAVE) {
    const float v = x[j];
    const uint16_t b = f2bf(v);
    if (Xb) Xb[size_t(row) * D + j] = b;
    const float w = bf2f(b);
    not_bf16 |= w != v;
    not_int |= !(rintf(v) == v && fabsf(v) <= 256.f);
    neg |= v < 0.f;
    big |= v > 255.f;
    s2 += double(v) * double(v);
    q2 += double(w) * double(w);
    e2 += (double(w) - double(v)) * (double(w) - double(v));
  }
  for (int o = 32; o >= 1; o >>= 1) {
    s2 += __shfl_xor(s2, o, TG_WAVE);
    q2 += __shfl_xor(q2, o, TG_WAVE);
    e2 += __shfl_xor(e2, o, TG_WAVE);
  }
  const uint64_t fb = __ballot(not_bf16 != 0), fi = __ballot(not_int != 0), fn = __ballot(neg != 0),
                 fg = __ballot(big != 0);
  if (threadIdx.x == 0) {
    nrm2[row] = float(s2);
    // norms rounded up, so the bound stays an upper bound
    nq[row] = float(sqrt(q2) * (1.0 + 1e-6));
    ne[row] = float(sqrt(e2) * (1.0 + 1e-6));
    if (fb) atomicOr(&flags[0], 1u);
    if (fi) atomicOr(&flags[1], 1u);
    if (fn) atomicOr(&flags[2], 1u);
    atomicMax(&flags[3], __float_as_uint(float(s2)));
    if (fg) atomicOr(&flags[4], 1u);
  }
}

// u8 copy of X, four values per thread (every value an integer in [0, 255], D % 4 == 0)
__global__ void k_make_u8(const float4* X, size_t n4, uint32_t* X8) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n4;
       i += size_t(gridDim.x) * blockDim.x) {
    const float4 v = X[i];
    X8[i] = uint32_t(v.x) | uint32_t(v.y) << 8 | uint32_t(v.z) << 16 | uint32_t(v.w) << 24;
  }
}

void launch_make_u8(const float* X, size_t n, uint8_t* X8, hipStream_t s) {
  const size_t n4 = n / 4;
  const uint32_t grid = uint32_t(std::min<size_t>((n4 + 255) / 256, 8192));
  k_make_u8<<<grid, 256, 0, s>>>(reinterpret_cast<const float4*>(X), n4,
                                 reinterpret_cast<uint32_t*>(X8));
}

// |x - 128|^2 of every u8 row, one wave per row (an integer below 2^24 for D <= 256, so
// exact as FP32)
__global__ void k_u8_shift_norms(const uint8_t* X8, uint32_t N, uint32_t D, float* n2s) {
  const uint32_t row = blockIdx.x;
  if (row >= N) return;
  const uint8_t* x = X8 + size_t(row) * D;
  int acc = 0;
  for (uint32_t j = threadIdx.x; j < D; j += TG_WAVE) {
    const int v = int(x[j]) - 128;
    acc += v * v;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, TG_WAVE);
  if (threadIdx.x == 0) n2s[row] = float(acc);
}

void launch_u8_shift_norms(const uint8_t* X8, uint32_t N, uint32_t D, float* n2s, hipStream_t s) {
  k_u8_shift_norms<<<N, 64, 0, s>>>(X8, N, D, n2s);
}

// FP8 e4m3 copy with a per-vector scale: q = e4m3(x / s), s = max|x| / 440 (below the
// format's 448 maximum), and the norms the screen bound needs: |s q| and |x - s q| in FP64,
// rounded up.  One wave per row, four elements per lane per step.
__global__ void k_screen_prep8(const float* X, uint32_t N, uint32_t D, uint8_t* Xq, float* sc,
                               float* nq, float* ne) {
  const uint32_t row = blockIdx.x;
  if (row >= N) return;
  const float* x = X + size_t(row) * D;
  float mx = 0.f;
  for (uint32_t j = threadIdx.x; j < D; j += TG_WAVE) mx = fmaxf(mx, fabsf(x[j]));
  for (int o = 32; o >= 1; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, TG_WAVE));
  const float s = mx > 0.f ? mx / 440.f : 1.f;
  double q2 = 0.0, e2 = 0.0;
  for (uint32_t j = 4 * threadIdx.x; j < D; j += 4 * TG_WAVE) {
    const float v0 = x[j], v1 = x[j + 1], v2 = x[j + 2], v3 = x[j + 3];
    int p = __builtin_amdgcn_cvt_pk_fp8_f32(v0 / s, v1 / s, 0, false);
    p = __builtin_amdgcn_cvt_pk_fp8_f32(v2 / s, v3 / s, p, true);
    *reinterpret_cast<int*>(Xq + size_t(row) * D + j) = p;
    typedef float f2 __attribute__((ext_vector_type(2)));
    const f2 lo = __builtin_amdgcn_cvt_pk_f32_fp8(p, false), hi = __builtin_amdgcn_cvt_pk_f32_fp8(p, true);
    const float v[4] = {v0, v1, v2, v3}, dq[4] = {lo.x, lo.y, hi.x, hi.y};
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const double w = double(s) * double(dq[b]);
      q2 += w * w;
      e2 += (double(v[b]) - w) * (double(v[b]) - w);
    }
  }
  for (int o = 32; o >= 1; o >>= 1) {
    q2 += __shfl_xor(q2, o, TG_WAVE);
    e2 += __shfl_xor(e2, o, TG_WAVE);
  }
  if (threadIdx.x == 0) {
    sc[row] = s;
    nq[row] = float(sqrt(q2) * (1.0 + 1e-6));
    ne[row] = float(sqrt(e2) * (1.0 + 1e-6));
  }
}

void launch_screen_prep8(const float* X, uint32_t N, uint32_t D, uint8_t* Xq, float* sc,
                         float* nq, float* ne, hipStream_t s) {
  k_screen_prep8<<<N, TG_WAVE, 0, s>>>(X, N, D, Xq, sc, nq, ne);
}

void launch_screen_prep(const float* X, uint32_t N, uint32_t D, uint16_t* Xb, float* nrm2,
                        float* nq, float* ne, uint32_t* flags, hipStream_t s) {
  k_screen_prep<<<N, TG_WAVE, 0, s>>>(X, N, D, Xb, nrm2, nq, ne, flags);
}

__device__ __forceinline__ float screen_bound(float na, float qa, float ea, float nb, float qb,
                                              float eb, float eta) {
  return 2.0f * (ea * qb + qa * eb + ea * eb) + eta * (na + nb + 2.0f * qa * qb);
}

// Products of the screen copies of the node's sample rows, staged through LDS DC elements
// at a time: wave w gets accB[c] = A rows 16w.. x B rows 16c.. and accA[c] = A rows 16w.. x
// A rows 16c.. (mfma layout: lane l, register r holds row 4(l/16)+r, column l%16 of the
// 16x16 block).  Rows past nA / nB are zero.
template <int PREC, int DC>
__device__ __forceinline__ void screen_products(const KnnArgs& a, const uint32_t* A,
                                                const uint32_t* B, int nA, int nB,
                                                unsigned char* sA, unsigned char* sB,
                                                f32x4 (&accB)[4], f32x4 (&accA)[4]) {
  constexpr int ESZ = PREC == 3 ? 1 : (PREC == 2 ? 2 : 4);  // element bytes
  constexpr int CPR = DC * ESZ / 16;                // 16-byte loads per staged row
  constexpr int RS = DC * ESZ + 16;                 // padded row stride in bytes
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  const int D = a.D;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    accB[c] = f32x4{0.f, 0.f, 0.f, 0.f};
    accA[c] = f32x4{0.f, 0.f, 0.f, 0.f};
  }
  const unsigned char* src = PREC == 3   ? reinterpret_cast<const unsigned char*>(a.Xq8)
                             : PREC == 2 ? reinterpret_cast<const unsigned char*>(a.Xb)
                                         : reinterpret_cast<const unsigned char*>(a.X);
  for (int d0 = 0; d0 < D; d0 += DC) {
    // stage rows [d0, d0+DC) of A and B: 16 B per thread-load, CPR loads per row
    for (int e = tid; e < 2 * 64 * CPR; e += 256) {
      const int op = e / (64 * CPR), r = (e / CPR) % 64, c16 = e % CPR;
      const int n = op == 0 ? nA : nB;
      uint4 v = make_uint4(0, 0, 0, 0);
      const int col = d0 + c16 * (16 / ESZ);
      if (r < n && col < D) {
        const uint32_t id = op == 0 ? A[r] : B[r];
        v = *reinterpret_cast<const uint4*>(src + (size_t(id) * D + col) * ESZ);
      }
      *reinterpret_cast<uint4*>((op == 0 ? sA : sB) + r * RS + c16 * 16) = v;
    }
    __syncthreads();
    if (PREC == 3) {
#pragma unroll
      for (int k0 = 0; k0 < DC; k0 += 32) {
        const int kk = k0 + 8 * (l >> 4);
        const long fa = *reinterpret_cast<const long*>(sA + (16 * w + (l & 15)) * RS + kk);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const long fb = *reinterpret_cast<const long*>(sB + (16 * c + (l & 15)) * RS + kk);
          const long fc = *reinterpret_cast<const long*>(sA + (16 * c + (l & 15)) * RS + kk);
          accB[c] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(fa, fb, accB[c], 0, 0, 0);
          accA[c] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(fa, fc, accA[c], 0, 0, 0);
        }
      }
    } else if (PREC == 2) {
#pragma unroll
      for (int k0 = 0; k0 < DC; k0 += 32) {
        const int kk = k0 + 8 * (l >> 4);
        const bf16x8 fa = *reinterpret_cast<const bf16x8*>(sA + (16 * w + (l & 15)) * RS + kk * 2);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const bf16x8 fb = *reinterpret_cast<const bf16x8*>(sB + (16 * c + (l & 15)) * RS + kk * 2);
          const bf16x8 fc = *reinterpret_cast<const bf16x8*>(sA + (16 * c + (l & 15)) * RS + kk * 2);
          accB[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa, fb, accB[c], 0, 0, 0);
          accA[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa, fc, accA[c], 0, 0, 0);
        }
      }
    } else {
#pragma unroll 4
      for (int k0 = 0; k0 < DC; k0 += 4) {
        const int kk = k0 + (l >> 4);
        const float fa = *reinterpret_cast<const float*>(sA + (16 * w + (l & 15)) * RS + kk * 4);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const float fb = *reinterpret_cast<const float*>(sB + (16 * c + (l & 15)) * RS + kk * 4);
          const float fc = *reinterpret_cast<const float*>(sA + (16 * c + (l & 15)) * RS + kk * 4);
          accB[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa, fb, accB[c], 0, 0, 0);
          accA[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa, fc, accA[c], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
}

// -------------------------------------------------------------------------------------
// Screened local join.  PREC 1: FP32 MFMA (16x16x4), PREC 2: BF16 MFMA (16x16x32),
// PREC 3: FP8 e4m3 MFMA (16x16x32) with per-vector scales applied to the product.
// 256 threads: wave w owns row block w (16 rows of A) against the 4 column blocks of B and
// of A.  Staging chunk: DC elements per row (128 BF16 or 64 FP32 = 256 B).
// -------------------------------------------------------------------------------------
template <int PREC, int MC, class T, int DC>
__global__ __launch_bounds__(256) void k_localjoin_screen(KnnArgs a) {
  // DC: elements per staged chunk (the tile's k depth per LDS round)
  constexpr int ESZ = PREC == 3 ? 1 : (PREC == 2 ? 2 : 4);  // element bytes
  constexpr int CPR = DC * ESZ / 16;                // 16-byte loads per staged row
  constexpr int RS = DC * ESZ + 16;                 // padded row stride in bytes
  constexpr int STAGE = 64 * RS;                    // one operand
  constexpr int MROW = 65;                          // padded distance-matrix row (floats)
  constexpr int SMEM = 2 * STAGE > 2 * 64 * MROW * 4 ? 2 * STAGE : 2 * 64 * MROW * 4;
  static_assert(DC % (PREC >= 2 ? 32 : 4) == 0 && CPR >= 1, "tile depth");
  // the distance matrices alias the staging area once the products are done
  __shared__ __attribute__((aligned(16))) unsigned char smem[SMEM];
  __shared__ uint32_t A[64], B[64];
  __shared__ float naA[64], qA[64], eA[64], wA[64];
  __shared__ float naB[64], qB[64], eB[64], wB[64];
  __shared__ float scA[64], scB[64];
  unsigned char* sA = smem;
  unsigned char* sB = smem + STAGE;
  float* MAB = reinterpret_cast<float*>(smem);
  float* MAA = MAB + 64 * MROW;

  const uint32_t u = row_of_block(a.perm, a.N);
  if (u >= a.N) return;
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  const int K = a.K, D = a.D;
  const int nA = a.na[u], nB = a.nb[u];
  if (tid < 64) {
    if (tid < nA) {
      const uint32_t v = a.H[size_t(u) * HROW_S + tid];
      A[tid] = v;
      naA[tid] = a.nrm2[v];
      qA[tid] = a.nq[v];
      eA[tid] = a.ne[v];
      wA[tid] = a.GD[size_t(v) * K + K - 1];
      if (PREC == 3) scA[tid] = a.sc8[v];
    }
  } else if (tid < 128) {
    const int j = tid - 64;
    if (j < nB) {
      const uint32_t v = a.H[size_t(u) * HROW_S + 64 + j];
      B[j] = v;
      naB[j] = a.nrm2[v];
      qB[j] = a.nq[v];
      eB[j] = a.ne[v];
      wB[j] = a.GD[size_t(v) * K + K - 1];
      if (PREC == 3) scB[j] = a.sc8[v];
    }
  }
  __syncthreads();

  f32x4 accB[4], accA[4];
  screen_products<PREC, DC>(a, A, B, nA, nB, sA, sB, accB, accA);
  // screen distances into the (aliased) matrices: C[i][j], i = 4*(l/16)+r, j = l%16
  const float INF = __uint_as_float(0x7F800000u);
#pragma unroll
  for (int c = 0; c < 4; ++c) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = 16 * w + 4 * (l >> 4) + r, j = 16 * c + (l & 15);
      float sb = INF, sa = INF;
      float pb = accB[c][r], pa = accA[c][r];
      if (PREC == 3) {
        // scale back: (s_i s_j) * sum q_i q_j
        pb *= (i < nA && j < nB) ? scA[i] * scB[j] : 0.f;
        pa *= (i < nA && j < nA) ? scA[i] * scA[j] : 0.f;
      }
      if (i < nA && j < nB && A[i] != B[j]) sb = (naA[i] + naB[j]) - 2.0f * pb;
      if (i < nA && j < nA && i != j) sa = (naA[i] + naA[j]) - 2.0f * pa;
      MAB[i * MROW + j] = sb;
      MAA[i * MROW + j] = sa;
    }
  }
  __syncthreads();

  // per (row or column, tile): the exact (d, id) minimum among the candidates whose bound
  // interval reaches below min(worst, min upper bound); pushed when d < worst
  // Jobs = (destination row, 16-wide candidate tile) of the three kinds (A rows x B tiles,
  // B columns x A tiles, A rows x A tiles), processed in chunks of QJ: one thread per job
  // computes the band mask from the screen matrix, the band candidates are queued, the 16
  // lane groups re-score the queue four pairs at a time, and each job keeps the exact
  // (d, id) minimum.  Candidates already in the destination's list would be dropped by the
  // merge and are not pushed.
  constexpr int QJ = 64;
  __shared__ uint32_t qent[QJ * LJ_TILE];  // (job in chunk << 8) | candidate column
  __shared__ float qdist[QJ * LJ_TILE];
  __shared__ int qoff[QJ + 1];
  const int gg = tid >> 4, t = tid & 15;
  const float eta = a.eta;
  const int tB = (nB + LJ_TILE - 1) / LJ_TILE, tA = (nA + LJ_TILE - 1) / LJ_TILE;
  // quotients by the tile counts (1..4) as multiply-shifts: exact for numerators < 21845
  const uint32_t mB = (65536u + uint32_t(tB) - 1) / uint32_t(max(tB, 1));
  const uint32_t mA = (65536u + uint32_t(tA) - 1) / uint32_t(max(tA, 1));
  auto divB = [&](int x) { return int((uint32_t(x) * mB) >> 16); };
  auto divA = [&](int x) { return int((uint32_t(x) * mA) >> 16); };
  const int nj0 = nA * tB, nj1 = nB * tA, njobs = nj0 + nj1 + nA * tA;
  uint32_t n_exact = 0;
  // the band masks of QJ jobs at a time are formed by wave 0 (exact screens take
  // k_localjoin_exact_screen)
  for (int base = 0; base < njobs; base += QJ) {
    int kind = 0, r = 0, lo = 0;
    uint32_t mask = 0;
    const int J = base + tid;
    if (tid < QJ && J < njobs) {
      if (J < nj0) {
        kind = 0; r = divB(J); lo = (J - r * tB) * LJ_TILE;
      } else if (J < nj0 + nj1) {
        kind = 1; r = divA(J - nj0); lo = (J - nj0 - r * tA) * LJ_TILE;
      } else {
        kind = 2; r = divA(J - nj0 - nj1); lo = (J - nj0 - nj1 - r * tA) * LJ_TILE;
      }
      const bool col = kind == 1, cb = kind == 0;
      const int nc = cb ? nB : nA, hi = min(nc, lo + LJ_TILE);
      const float* M = kind == 2 ? MAA : MAB;
      const float worst = col ? wB[r] : wA[r];
      const float nd = col ? naB[r] : naA[r], qd = col ? qB[r] : qA[r];
      const float ed = col ? eB[r] : eA[r];
      float sv[LJ_TILE], ev[LJ_TILE];
      float tau = worst;
#pragma unroll
      for (int q = 0; q < LJ_TILE; ++q) {
        const int j = lo + q;
        sv[q] = INF;
        ev[q] = 0.f;
        if (j < hi) {
          sv[q] = col ? M[j * MROW + r] : M[r * MROW + j];
          if (sv[q] < INF) {
            ev[q] = screen_bound(nd, qd, ed, cb ? naB[j] : naA[j], cb ? qB[j] : qA[j],
                                 cb ? eB[j] : eA[j], eta);
            tau = fminf(tau, sv[q] + ev[q]);
          }
        }
      }
#pragma unroll
      for (int q = 0; q < LJ_TILE; ++q)
        if (sv[q] < INF && !(sv[q] - ev[q] > tau)) mask |= 1u << q;
    }
    // ---- queue offsets (wave 0 scan)
    if (tid < QJ) {
      const int c = __popc(mask);
      const int incl = wave_incl_scan(c);
      qoff[tid + 1] = incl;
      if (tid == 0) qoff[0] = 0;
      int pos = incl - c;
      uint32_t m = mask;
      while (m) {
        const int q = __ffs(m) - 1;
        m &= m - 1;
        qent[pos++] = (uint32_t(tid) << 8) | uint32_t(lo + q) | (uint32_t(kind) << 6);
      }
    }
    __syncthreads();
    const int nq = qoff[QJ];
    // ---- exact re-scores, four pairs in flight per 16-lane group
    for (int e0 = 0; e0 < nq; e0 += 64) {
      const T* Xr = a.Xr16 ? reinterpret_cast<const T*>(a.Xr16) : reinterpret_cast<const T*>(a.X);
      const T* ra[4];
      const T* rb[4];
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int e = e0 + gg + 16 * b;
        ra[b] = rb[b] = Xr;  // unused slots read row 0
        if (e < nq) {
          const uint32_t q = qent[e];
          const int jk = (q >> 6) & 3, jc = q & 63;
          const int J2 = base + int(q >> 8);
          const int rr = jk == 0 ? divB(J2) : (jk == 1 ? divA(J2 - nj0) : divA(J2 - nj0 - nj1));
          const uint32_t dest = jk == 1 ? B[rr] : A[rr];
          const uint32_t cid = jk == 0 ? B[jc] : A[jc];
          ra[b] = Xr + size_t(dest) * D;
          rb[b] = Xr + size_t(cid) * D;
        }
      }
      float dd[4];
      dist16_pairs<MC, 4>(ra, rb, D, t, dd);
      if (t == 0) {
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const int e = e0 + gg + 16 * b;
          if (e < nq) qdist[e] = dd[b];
        }
      }
    }
    n_exact += nq;
    __syncthreads();
    // ---- per job: exact minimum of the band, push
    if (tid < QJ && mask) {
      const uint32_t dest = kind == 1 ? B[r] : A[r];
      const float worst = kind == 1 ? wB[r] : wA[r];
      uint64_t best = TG_KEY_INF;
      for (int e = qoff[tid]; e < qoff[tid + 1]; ++e) {
        const int jc = qent[e] & 63;
        best = min(best, mk_key(qdist[e], kind == 0 ? B[jc] : A[jc]));
      }
      if (key_d(best) < worst && !row_has_key(a.G, a.GD, K, dest, best)) pb_push(a.pc, dest, best);
    }
    __syncthreads();
  }
  if (a.counters && tid == 0) {
    // [0] pair evaluations (screen), [1] exact re-scores
    atomicAdd(&cnt_slot(a.counters)[0], (unsigned long long)(nA * nB + nA * (nA - 1) / 2));
    atomicAdd(&cnt_slot(a.counters)[1], (unsigned long long)n_exact);
  }
}

// Local join when the screen distances are the canonical distances (integer data, FP32 or
// BF16 products): each job's answer is the (s, id) minimum of its 16-wide tile, taken from
// the accumulators by cross-lane minima instead of distance matrices in LDS, so the kernel
// holds only the staging area (and runs more workgroups per CU).  Jobs, as in the screened
// kernel: A rows x B tiles, B columns x A tiles, A rows x A tiles.  Lane t of group g owns
// row job (row 16w + 4g + (t & 3), tile t >> 2) of both row kinds, and lane l owns column
// job (column l, tile w).
// (d, id) minimum over the lanes l ^ x for the masks x in [x0, x1) (powers of two): the
// smallest d, then the smallest id among the lanes holding it.  The exact screen's distances
// are integers >= 0 or +inf, whose FP32 bits order as unsigned integers, so both minima are
// integer minima (no NaN handling), and the result is the minimum of the keys mk_key(d, id).
template <int X0, int X1>
__device__ __forceinline__ uint64_t key_min_lanes(float d, uint32_t id) {
  const uint32_t db = __float_as_uint(d);
  uint32_t m = db;
#pragma unroll
  for (int x = X0; x < X1; x <<= 1) m = min(m, lane_xor32(m, x));
  uint32_t k = db == m ? id : 0xFFFFFFFFu;
#pragma unroll
  for (int x = X0; x < X1; x <<= 1) k = min(k, lane_xor32(k, x));
  return (uint64_t(m) << 32) | k;
}

// Minimum over the 16 lanes of the lane's DPP row, in every lane of it: quad exchanges, then
// row rotations by 4 and 8 (a minimum needs no exact partners).  Each step is one DPP-fused
// v_min_u32.
template <int CTRL>
__device__ __forceinline__ uint32_t min_dpp(uint32_t m) {
  return min(m, uint32_t(__builtin_amdgcn_update_dpp(int(0xFFFFFFFFu), int(m), CTRL, 0xF, 0xF,
                                                     false)));
}
__device__ __forceinline__ uint32_t row_min16(uint32_t m) {
  m = min_dpp<0xB1>(m);   // quad: lane ^ 1
  m = min_dpp<0x4E>(m);   // quad: lane ^ 2
  m = min_dpp<0x124>(m);  // row_ror:4
  return min_dpp<0x128>(m);  // row_ror:8
}
// key_min_lanes<1, 16> with row_min16: the (d, id) minimum over the lane's 16-lane row.
__device__ __forceinline__ uint64_t key_min_row16(float d, uint32_t id) {
  const uint32_t db = __float_as_uint(d);
  const uint32_t m = row_min16(db);
  return (uint64_t(m) << 32) | row_min16(db == m ? id : 0xFFFFFFFFu);
}

// Membership of Q keys in kNN rows (rows sorted by (distance, id), canonical distances):
// found[q] = key[q] is an entry of row v[q], for the searches with on[q].  The binary
// searches run in lockstep, so the dependent loads of one step are in flight together.
template <int Q>
__device__ __forceinline__ void rows_have_keys(const uint32_t* G, const float* GD, int K,
                                               const uint32_t (&v)[Q], const uint64_t (&key)[Q],
                                               const bool (&on)[Q], bool (&found)[Q]) {
  int lo[Q], hi[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    lo[q] = 0;
    hi[q] = on[q] ? K : 0;
  }
  for (;;) {
    bool any = false;
#pragma unroll
    for (int q = 0; q < Q; ++q) any = any || lo[q] < hi[q];
    if (!any) break;
    uint64_t kk[Q];
    int mid[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      mid[q] = (lo[q] + hi[q]) >> 1;
      kk[q] = lo[q] < hi[q] ? mk_key(GD[size_t(v[q]) * K + mid[q]], G[size_t(v[q]) * K + mid[q]])
                            : 0ull;
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      if (lo[q] < hi[q]) {
        if (kk[q] < key[q]) lo[q] = mid[q] + 1;
        else hi[q] = mid[q];
      }
    }
  }
#pragma unroll
  for (int q = 0; q < Q; ++q)
    found[q] = on[q] && lo[q] < K &&
               mk_key(GD[size_t(v[q]) * K + lo[q]], G[size_t(v[q]) * K + lo[q]]) == key[q];
}

// Column block c of an exact screen (wave w, lane l = 16g + t): accB holds A rows
// 16w + 4g + r against B row 16c + t, accA the same A rows against A row 16c + t, as exact
// products with s = (na[i] + na[j]) - 2 acc.  Folds the block into the row jobs' minima
// (row_b / row_a: the lane t == 4c + r owns job (16w + 4g + r, c)) and into this lane's
// column job of the block (col_b: the minimum over the wave's 16 A rows, held by g == c).
__device__ __forceinline__ void lj_block_minima(int c, const f32x4& accB, const f32x4& accA,
                                                const uint32_t* A, const uint32_t* B,
                                                const float* naA, const float* naB, int nA,
                                                int nB, int w, int g, int t, uint64_t& row_b,
                                                uint64_t& row_a, uint64_t& col_b) {
  const float INF = __uint_as_float(0x7F800000u);
  const int j = 16 * c + t;
  float cd = INF;  // column job: this lane's 4 rows, then across the groups
  uint32_t cid = 0xFFFFFFFFu;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = 16 * w + 4 * g + r;
    const float sb = (i < nA && j < nB && A[i] != B[j]) ? (naA[i] + naB[j]) - 2.0f * accB[r] : INF;
    const float sa = (i < nA && j < nA && i != j) ? (naA[i] + naA[j]) - 2.0f * accA[r] : INF;
    const uint64_t kb = key_min_row16(sb, B[j]);
    const uint64_t ka = key_min_row16(sa, A[j]);
    if (t == 4 * c + r) {
      row_b = kb;
      row_a = ka;
    }
    if (sb < cd || (sb == cd && A[i] < cid)) {
      cd = sb;
      cid = A[i];
    }
  }
  const uint64_t kc = key_min_lanes<16, 64>(cd, cid);
  if (g == c) col_b = kc;
}

// The jobs' pushes: a job's minimum enters when it beats the destination's worst entry and
// is not already listed (the merge would drop it); the three row searches run together.
// Keys for one destination take one counter update and consecutive slots, so they share
// cache lines: the row jobs of A row i sit in the four lanes t = (i % 4) + 4c of a 16-lane
// row (up to 8 keys), the column jobs of B row l in lane l of the four waves (up to 4 keys,
// gathered through colk, 4 x 64 keys of LDS that no wave reads any more).  The slot order
// of a list is immaterial (its reduction sorts).
__device__ __forceinline__ void lj_push(const KnnArgs& a, const uint32_t* A, const uint32_t* B,
                                        const float* wA, const float* wB, int nA, int nB,
                                        int w, int g, int t, int l, uint64_t row_b,
                                        uint64_t row_a, uint64_t col_b, uint64_t* colk) {
  const int i = 16 * w + 4 * g + (t & 3);
  const bool inA = i < nA, inB = l < nB;
  const uint32_t destA = inA ? A[i] : 0u, destB = inB ? B[l] : 0u;
  const float worstA = inA ? wA[i] : 0.f, worstB = inB ? wB[l] : 0.f;
  const uint32_t dv[3] = {destA, destA, destB};
  const uint64_t kv[3] = {row_b, row_a, col_b};
  const bool on[3] = {inA && key_d(row_b) < worstA, inA && key_d(row_a) < worstA,
                      inB && key_d(col_b) < worstB};
  bool found[3];
  rows_have_keys<3>(a.G, a.GD, a.K, dv, kv, on, found);
  // row jobs: counts of the four lanes c = t / 4, their prefix in c order, one counter update
  const bool pb = on[0] && !found[0], pa = on[1] && !found[1];
  const int cnt = int(pb) + int(pa), c = t >> 2;
  const int x4 = int(lane_xor32(uint32_t(cnt), 4));
  const int p1 = cnt + x4;
  const int x8 = int(lane_xor32(uint32_t(p1), 8));
  const int total = p1 + x8;
  const int prefix = ((c & 1) ? x4 : 0) + ((c & 2) ? x8 : 0);
  uint32_t base = 0;
  if (c == 0 && total > 0) base = atomicAdd(&a.pc.cnt[destA], uint32_t(total));
  base = uint32_t(__shfl(int(base), (l & ~15) | (t & 3), TG_WAVE));
  if (pb) pb_put(a.pc, destA, base + uint32_t(prefix), row_b);
  if (pa) pb_put(a.pc, destA, base + uint32_t(prefix) + (pb ? 1u : 0u), row_a);
  // column jobs: wave 0 pushes the four waves' keys of B row l together
  __syncthreads();
  colk[w * 64 + l] = (on[2] && !found[2]) ? col_b : TG_KEY_INF;
  __syncthreads();
  if (w == 0 && inB) {
    uint64_t k4[4];
    uint32_t n = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      k4[q] = colk[q * 64 + l];
      n += k4[q] != TG_KEY_INF ? 1u : 0u;
    }
    if (n) {
      uint32_t s = atomicAdd(&a.pc.cnt[destB], n);
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (k4[q] != TG_KEY_INF) pb_put(a.pc, destB, s++, k4[q]);
    }
  }
}

// The node's sample lists, their norms (na) and their destinations' worst distances (wl).
__device__ __forceinline__ void lj_lists(const KnnArgs& a, uint32_t u, const float* norms,
                                         int nA, int nB, uint32_t* A, uint32_t* B, float* naA,
                                         float* wA, float* naB, float* wB) {
  const int tid = threadIdx.x, K = a.K;
  if (tid < 64) {
    const bool in = tid < nA;
    const uint32_t v = in ? a.H[size_t(u) * HROW_S + tid] : 0u;
    A[tid] = v;
    naA[tid] = in ? norms[v] : 0.f;
    wA[tid] = in ? a.GD[size_t(v) * K + K - 1] : 0.f;
  } else if (tid < 128) {
    const int j = tid - 64;
    const bool in = j < nB;
    const uint32_t v = in ? a.H[size_t(u) * HROW_S + 64 + j] : 0u;
    B[j] = v;
    naB[j] = in ? norms[v] : 0.f;
    wB[j] = in ? a.GD[size_t(v) * K + K - 1] : 0.f;
  }
}

template <int PREC, int DC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
void k_localjoin_exact_screen(KnnArgs a) {
  constexpr int ESZ = PREC == 2 ? 2 : 4;
  constexpr int STAGE = 64 * (DC * ESZ + 16);
  static_assert(PREC == 1 || PREC == 2, "exact screens are FP32 or BF16 products");
  __shared__ __attribute__((aligned(16))) unsigned char smem[2 * STAGE];
  __shared__ uint32_t A[64], B[64];
  __shared__ float naA[64], wA[64], naB[64], wB[64];
  const uint32_t u = row_of_block(a.perm, a.N);
  if (u >= a.N) return;
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, g = l >> 4, t = l & 15;
  const int nA = a.na[u], nB = a.nb[u];
  lj_lists(a, u, a.nrm2, nA, nB, A, B, naA, wA, naB, wB);
  __syncthreads();
  f32x4 accB[4], accA[4];
  screen_products<PREC, DC>(a, A, B, nA, nB, smem, smem + STAGE, accB, accA);
  uint64_t row_b = TG_KEY_INF, row_a = TG_KEY_INF, col_b = TG_KEY_INF;
#pragma unroll
  for (int c = 0; c < 4; ++c)
    lj_block_minima(c, accB[c], accA[c], A, B, naA, naB, nA, nB, w, g, t, row_b, row_a, col_b);
  lj_push(a, A, B, wA, wB, nA, nB, w, g, t, l, row_b, row_a, col_b,
          reinterpret_cast<uint64_t*>(smem));
  if (a.counters && tid == 0)
    atomicAdd(&cnt_slot(a.counters)[0], (unsigned long long)(nA * nB + nA * (nA - 1) / 2));
}

typedef int i32x4 __attribute__((ext_vector_type(4)));

// Exact screen on int8 MFMA for u8 data (every value an integer in [0, 255], D <= DC <= 256):
// rows are staged whole as x - 128 (the byte XOR 0x80, a signed int8), and
// v_mfma_i32_16x16x64_i8 gives a'.b' exactly, so s = |a'|^2 + |b'|^2 - 2 a'.b' = |a - b|^2
// exactly (|a'|^2 from n2s): the canonical distance of this data, as in the BF16 exact
// screen.  Half the staged bytes and MFMA instructions of the BF16 rows; the column blocks
// are formed and folded one at a time, so only two accumulators are live.
template <int DC>
__global__ __launch_bounds__(256) void k_localjoin_u8_screen(KnnArgs a) {
  constexpr int RS = DC + 16;  // padded LDS row stride in bytes
  constexpr int CPR = DC / 16;  // 16-byte pieces per staged row
  constexpr int KQ = DC / 64;   // MFMA K-steps
  __shared__ __attribute__((aligned(16))) unsigned char sA[64 * RS], sB[64 * RS];
  __shared__ uint32_t A[64], B[64];
  __shared__ float naA[64], wA[64], naB[64], wB[64];
  const uint32_t u = row_of_block(a.perm, a.N);
  if (u >= a.N) return;
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, g = l >> 4, t = l & 15;
  const int D = a.D;
  const int nA = a.na[u], nB = a.nb[u];
  lj_lists(a, u, a.n2s, nA, nB, A, B, naA, wA, naB, wB);
  __syncthreads();
  // stage both lists' rows: bytes past D and rows past the list ends are 0 (no product)
  for (int e = tid; e < 2 * 64 * CPR; e += 256) {
    const int op = e / (64 * CPR), r = (e / CPR) % 64, col = (e % CPR) * 16;
    const int n = op == 0 ? nA : nB;
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    if (r < n && col < D) {
      const uint8_t* src = a.X8 + size_t(op == 0 ? A[r] : B[r]) * D + col;
      if ((D & 15) == 0) {
        v = *reinterpret_cast<const uint4*>(src);
      } else {
        v.x = *reinterpret_cast<const uint32_t*>(src);
        if (col + 4 < D) v.y = *reinterpret_cast<const uint32_t*>(src + 4);
        if (col + 8 < D) v.z = *reinterpret_cast<const uint32_t*>(src + 8);
        if (col + 12 < D) v.w = *reinterpret_cast<const uint32_t*>(src + 12);
      }
      v.x ^= col < D ? 0x80808080u : 0u;
      v.y ^= col + 4 < D ? 0x80808080u : 0u;
      v.z ^= col + 8 < D ? 0x80808080u : 0u;
      v.w ^= col + 12 < D ? 0x80808080u : 0u;
    }
    *reinterpret_cast<uint4*>((op == 0 ? sA : sB) + r * RS + col) = v;
  }
  __syncthreads();
  // A-operand pieces of this wave's 16 rows, every K-step (lane group g: bytes 16g..16g+15)
  i32x4 fa[KQ];
#pragma unroll
  for (int k = 0; k < KQ; ++k)
    fa[k] = *reinterpret_cast<const i32x4*>(sA + (16 * w + t) * RS + 64 * k + 16 * g);
  uint64_t row_b = TG_KEY_INF, row_a = TG_KEY_INF, col_b = TG_KEY_INF;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    i32x4 pB = {0, 0, 0, 0}, pA = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < KQ; ++k) {
      const i32x4 fb = *reinterpret_cast<const i32x4*>(sB + (16 * c + t) * RS + 64 * k + 16 * g);
      const i32x4 fc = *reinterpret_cast<const i32x4*>(sA + (16 * c + t) * RS + 64 * k + 16 * g);
      pB = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[k], fb, pB, 0, 0, 0);
      pA = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[k], fc, pA, 0, 0, 0);
    }
    const f32x4 accB = {float(pB[0]), float(pB[1]), float(pB[2]), float(pB[3])};
    const f32x4 accA = {float(pA[0]), float(pA[1]), float(pA[2]), float(pA[3])};
    lj_block_minima(c, accB, accA, A, B, naA, naB, nA, nB, w, g, t, row_b, row_a, col_b);
  }
  lj_push(a, A, B, wA, wB, nA, nB, w, g, t, l, row_b, row_a, col_b,
          reinterpret_cast<uint64_t*>(sA));
  if (a.counters && tid == 0)
    atomicAdd(&cnt_slot(a.counters)[0], (unsigned long long)(nA * nB + nA * (nA - 1) / 2));
}

// tile 0: 256-byte staged rows (256 FP8 / 128 BF16 / 64 FP32 elements per round); tile 1:
// 64-byte rows (64 / 32 / 16 elements per round): less zero padding at D not a multiple of
// the depth, more LDS rounds
void launch_knn_localjoin_screen(const KnnArgs& a, int precision, int tile, hipStream_t s) {
  if (a.exact_screen && precision == KP_BF16 && a.X8 && a.n2s && a.D <= 256) {
    // u8 data: the BF16 screen's products on int8 MFMA (exact either way)
    const dim3 grid(grid_rows(a.perm, a.N));
    if (a.D <= 128) k_localjoin_u8_screen<128><<<grid, 256, 0, s>>>(a);
    else k_localjoin_u8_screen<256><<<grid, 256, 0, s>>>(a);
    return;
  }
  if (a.exact_screen && (precision == KP_BF16 || precision == KP_FP32)) {
    // 256-byte staged rows (128 BF16 / 64 FP32 elements), 64-byte rows with tile 1
    const dim3 grid(grid_rows(a.perm, a.N));
    if (precision == KP_BF16) {
      if (tile == 1) k_localjoin_exact_screen<2, 32><<<grid, 256, 0, s>>>(a);
      else k_localjoin_exact_screen<2, 128><<<grid, 256, 0, s>>>(a);
    } else {
      if (tile == 1) k_localjoin_exact_screen<1, 16><<<grid, 256, 0, s>>>(a);
      else k_localjoin_exact_screen<1, 64><<<grid, 256, 0, s>>>(a);
    }
    return;
  }
  const int mc = (int(a.D) + 63) / 64;
#define LJS(P, M, T, DC) \
  k_localjoin_screen<P, M, T, DC><<<grid_rows(a.perm, a.N), 256, 0, s>>>(a)
#define LJS_DC(P, M, T)                                                    \
  do {                                                                     \
    if (tile == 1) LJS(P, M, T, (P == 3 ? 64 : P == 2 ? 32 : 16));        \
    else LJS(P, M, T, (P == 3 ? 256 : P == 2 ? 128 : 64));                 \
  } while (0)
#define LJS_MC(P, T)                                                       \
  do {                                                                     \
    if (mc <= 2) LJS_DC(P, 2, T); else if (mc <= 4) LJS_DC(P, 4, T); else LJS_DC(P, 16, T); \
  } while (0)
  if (precision == KP_FP8) {
    if (a.Xr16) LJS_MC(3, uint16_t); else LJS_MC(3, float);
  } else if (precision == KP_BF16) {
    if (a.Xr16) LJS_MC(2, uint16_t); else LJS_MC(2, float);
  } else {
    if (a.Xr16) LJS_MC(1, uint16_t); else LJS_MC(1, float);
  }
#undef LJS_MC
#undef LJS_DC
#undef LJS
}

}  // namespace tg
