// learn.hip -- kernels and host driver of the learned initial lists (see learn.hpp).
//
// GEMMs: one LDS-tiled kernel C = A * B^T with A [M x Kd] and B [Nc x Kd] both K-contiguous
// BF16 (the operand layout v_mfma_f32_16x16x32_bf16 reads directly), 64 x 64 tiles, four
// waves of 2 x 2 MFMA tiles, split-K over blockIdx.z with partials reduced in a fixed order.
// The backward pass keeps transposed copies (Xn^T, H^T, dZ^T, dH^T) so that every product,
// including the weight gradients, is the same NT form; a row of ones appended to Xn^T and
// H^T makes the bias gradients columns of the weight-gradient products.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "kernels.hpp"
#include "learn.hpp"

namespace tg {

namespace {

constexpr int GT = 64;  // GEMM tile edge
constexpr int GK = 32;  // GEMM k step
constexpr int GPAD = 8;
// Grouping chunk: at least 1024 vectors and at most about 1024 chunks, so the stable scatter
// (one wave per chunk, serial within it) has about four waves per CU and the per-chunk
// counts stay near 1024 x C words.
inline uint32_t grp_chunk(uint32_t N) {
  const uint32_t c = ((N + 1023) / 1024 + 63) / 64 * 64;
  return c < 1024 ? 1024 : c;
}
constexpr uint32_t MAX_BUCKETS = 4096;
constexpr uint32_t ENC_CHUNK = 131072;

enum { EPI_H = 0, EPI_Z = 1, EPI_DH = 2, EPI_PART = 3 };

struct GemmArgs {
  const uint16_t* A;
  const uint16_t* B;
  uint32_t M, Nc, Kd, lda, ldb;
  float* Cf;            // EPI_Z: [M x ldc]; EPI_PART: partial z at Cf + z * zstride
  uint32_t ldc;
  size_t zstride;
  uint16_t* Cb;         // EPI_H row-major BF16 output
  uint32_t ldcb;
  uint16_t* CT;         // EPI_H / EPI_DH transposed BF16 output [Nc x ldct]
  uint32_t ldct;
  const float* bias;    // bias of output column j at bias[j * bias_stride]
  uint32_t bias_stride;
  const uint16_t* mask; // EPI_DH: relu mask source H [M x ldm]
  uint32_t ldm;
};

template <int EPI>
__global__ __launch_bounds__(256) void k_gemm_nt(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) uint16_t As[GT][GK + GPAD];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[GT][GK + GPAD];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const uint32_t row0 = blockIdx.x * GT, col0 = blockIdx.y * GT;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int kt = int(g.Kd / GK);
  const int per = (kt + int(gridDim.z) - 1) / int(gridDim.z);
  const int kb0 = int(blockIdx.z) * per, kb1 = min(kt, kb0 + per);
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int lr = tid >> 2, lk = (tid & 3) * 8;
  for (int kb = kb0; kb < kb1; ++kb) {
    const uint32_t kk = uint32_t(kb) * GK + lk;
    uint4 va = make_uint4(0, 0, 0, 0), vb = make_uint4(0, 0, 0, 0);
    if (row0 + lr < g.M) va = *reinterpret_cast<const uint4*>(g.A + size_t(row0 + lr) * g.lda + kk);
    if (col0 + lr < g.Nc) vb = *reinterpret_cast<const uint4*>(g.B + size_t(col0 + lr) * g.ldb + kk);
    *reinterpret_cast<uint4*>(&As[lr][lk]) = va;
    *reinterpret_cast<uint4*>(&Bs[lr][lk]) = vb;
    __syncthreads();
    bf16x8 fa[2], fb[2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
      fa[i] = *reinterpret_cast<const bf16x8*>(&As[wr + 16 * i + (lane & 15)][8 * (lane >> 4)]);
#pragma unroll
    for (int j = 0; j < 2; ++j)
      fb[j] = *reinterpret_cast<const bf16x8*>(&Bs[wc + 16 * j + (lane & 15)][8 * (lane >> 4)]);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  // lane holds C[4 * (lane / 16) + r][lane % 16] of each 16 x 16 tile
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const uint32_t col = col0 + wc + 16 * j + (lane & 15);
      if (col >= g.Nc) continue;
      const uint32_t rb = row0 + wr + 16 * i + 4 * (lane >> 4);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const uint32_t row = rb + r;
        if (row >= g.M) continue;
        float v = acc[i][j][r];
        if (EPI == EPI_H) {
          v = fmaxf(v + g.bias[size_t(col) * g.bias_stride], 0.f);
          const uint16_t b = f2bf(v);
          g.Cb[size_t(row) * g.ldcb + col] = b;
          if (g.CT) g.CT[size_t(col) * g.ldct + row] = b;
        } else if (EPI == EPI_Z) {
          g.Cf[size_t(row) * g.ldc + col] = v + g.bias[size_t(col) * g.bias_stride];
        } else if (EPI == EPI_DH) {
          const float h = bf2f(g.mask[size_t(row) * g.ldm + col]);
          g.CT[size_t(col) * g.ldct + row] = f2bf(h > 0.f ? v : 0.f);
        } else {
          g.Cf[blockIdx.z * g.zstride + size_t(row) * g.ldc + col] = v;
        }
      }
    }
}

template <int EPI>
void gemm(const GemmArgs& g, uint32_t splits, hipStream_t st) {
  dim3 grid((g.M + GT - 1) / GT, (g.Nc + GT - 1) / GT, splits);
  k_gemm_nt<EPI><<<grid, 256, 0, st>>>(g);
}

// xn = bf16((x - mu) * s) for rows X[rows[m]] (or X[first + m]); Xn [M x D], XnT [D x ldt]
__global__ __launch_bounds__(256) void k_mlp_gather(const float* __restrict__ X, uint32_t D,
                                                    const uint32_t* __restrict__ rows,
                                                    uint32_t first, uint32_t M,
                                                    const float* __restrict__ mu,
                                                    const float* __restrict__ scal, uint16_t* Xn,
                                                    uint16_t* XnT, uint32_t ldt) {
  __shared__ uint16_t tile[64][65];
  const int tid = threadIdx.x;
  const uint32_t m0 = blockIdx.x * 64, d0 = blockIdx.y * 64;
  const float s = scal[0];
  for (int e = tid; e < 64 * 64; e += 256) {
    const int r = e >> 6, c = e & 63;
    const uint32_t m = m0 + r, d = d0 + c;
    if (m < M && d < D) {
      const uint32_t node = rows ? rows[m] : first + m;
      const uint16_t b = f2bf((X[size_t(node) * D + d] - mu[d]) * s);
      Xn[size_t(m) * D + d] = b;
      tile[r][c] = b;
    }
  }
  if (!XnT) return;
  __syncthreads();
  for (int e = tid; e < 64 * 64; e += 256) {
    const int c = e >> 6, r = e & 63;
    const uint32_t m = m0 + r, d = d0 + c;
    if (m < M && d < D) XnT[size_t(d) * ldt + m] = tile[r][c];
  }
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, TG_WAVE);
  return v;
}

// Contrastive loss and its gradient with respect to the codes: one workgroup of four waves
// per pair of anchors (B is even), so each negative's code is read once for both; thread tid
// takes the negatives c = tid, tid + 256, ... for each anchor in that order, the wave
// partials are added in wave order, and wave a finishes anchor b0 + a (deterministic, and
// the same sums as one anchor per workgroup).
constexpr int LOSS_AP = 2;
__global__ __launch_bounds__(256) void k_mlp_loss(const float* __restrict__ Z, uint32_t B,
                                                  uint32_t P, float margin, uint32_t M,
                                                  uint16_t* dZb, uint16_t* dZT, float* lossp) {
  __shared__ float red[LOSS_AP][4][MLP_CODE + 2];
  __shared__ float zas[LOSS_AP][MLP_CODE];
  const uint32_t b0 = blockIdx.x * LOSS_AP;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  if (tid < LOSS_AP * int(MLP_CODE)) {
    const uint32_t b = b0 + uint32_t(tid) / MLP_CODE;
    zas[tid / MLP_CODE][tid % MLP_CODE] = b < B ? Z[size_t(b) * MLP_CODE + tid % MLP_CODE] : 0.f;
  }
  __syncthreads();
  float acc[LOSS_AP][MLP_CODE];
  float nact[LOSS_AP], lneg[LOSS_AP];
#pragma unroll
  for (int q = 0; q < LOSS_AP; ++q) {
    nact[q] = lneg[q] = 0.f;
#pragma unroll
    for (int k = 0; k < int(MLP_CODE); ++k) acc[q][k] = 0.f;
  }
  for (uint32_t c = tid; c < B; c += 256) {
    const float* zc = Z + size_t(c) * MLP_CODE;
    float zv[MLP_CODE];
#pragma unroll
    for (int k = 0; k < int(MLP_CODE); ++k) zv[k] = zc[k];
#pragma unroll
    for (int q = 0; q < LOSS_AP; ++q) {
      const uint32_t b = b0 + uint32_t(q);
      if (c == b || b >= B) continue;
      float d2 = 0.f;
#pragma unroll
      for (int k = 0; k < int(MLP_CODE); ++k) {
        const float e = zas[q][k] - zv[k];
        d2 = fmaf(e, e, d2);
      }
      if (d2 < margin) {
        nact[q] += 1.f;
        lneg[q] += margin - d2;
#pragma unroll
        for (int k = 0; k < int(MLP_CODE); ++k) acc[q][k] += zv[k];
      }
    }
  }
#pragma unroll
  for (int q = 0; q < LOSS_AP; ++q) {
    const float n = wave_sum(nact[q]), l = wave_sum(lneg[q]);
#pragma unroll
    for (int k = 0; k < int(MLP_CODE); ++k) {
      const float v = wave_sum(acc[q][k]);
      if (lane == k) red[q][wave][2 + k] = v;
    }
    if (lane == 0) {
      red[q][wave][0] = n;
      red[q][wave][1] = l;
    }
  }
  __syncthreads();
  if (wave >= LOSS_AP) return;
  const int qa = wave;
  const uint32_t b = b0 + uint32_t(qa);
  if (b >= B) return;
  const int k = lane & 31;
  const float na = ((red[qa][0][0] + red[qa][1][0]) + red[qa][2][0]) + red[qa][3][0];
  const float ln = ((red[qa][0][1] + red[qa][1][1]) + red[qa][2][1]) + red[qa][3][1];
  const float accs =
      ((red[qa][0][2 + k] + red[qa][1][2 + k]) + red[qa][2][2 + k]) + red[qa][3][2 + k];
  const float cpos = 2.f / float(B * P), cneg = 4.f / (float(B) * float(B - 1));
  const float zak = zas[qa][k];
  float sp = 0.f, lpos = 0.f;
  for (uint32_t j = 0; j < P; ++j) {
    const uint32_t row = B + b * P + j;
    const float zp = Z[size_t(row) * MLP_CODE + k];
    const float e = zak - zp;
    sp += e;
    lpos = fmaf(e, e, lpos);
    if (lane < int(MLP_CODE)) {
      const uint16_t g = f2bf(cpos * (zp - zak));
      dZb[size_t(row) * MLP_CODE + k] = g;
      dZT[size_t(k) * M + row] = g;
    }
  }
  if (lane < int(MLP_CODE)) {
    const uint16_t g = f2bf(cpos * sp - cneg * (na * zak - accs));
    dZb[size_t(b) * MLP_CODE + k] = g;
    dZT[size_t(k) * M + b] = g;
  }
  lpos = wave_sum(lane < int(MLP_CODE) ? lpos : 0.f);
  if (lane == 0) lossp[b] = lpos / float(B * P) + ln / (float(B) * float(B - 1));
}

// Adam over the flat parameter block (and the BF16 weight copies)
__global__ __launch_bounds__(256) void k_mlp_adam(float* prm, float* mom, float* vel,
                                                  const float* __restrict__ gpart, uint32_t D,
                                                  uint32_t splits, int mode, float lr,
                                                  float b1, float b2, float eps, float bc1,
                                                  float bc2, uint16_t* W1b, uint16_t* W2b,
                                                  uint16_t* W2Tb, float* grad_out) {
  const uint32_t n1 = mlp_n1(D), n2 = mlp_n2();
  const uint32_t idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= n1 + n2) return;
  float p = prm[idx];
  if (mode >= 1) {
    float g = 0.f;
    if (idx < n1) {
      for (uint32_t z = 0; z < splits; ++z) g += gpart[size_t(z) * n1 + idx];
    } else {
      const uint32_t r = idx - n1;
      for (uint32_t z = 0; z < splits; ++z) g += gpart[size_t(splits) * n1 + size_t(z) * n2 + r];
    }
    if (mode == 1) {
      grad_out[idx] = g;
      return;
    }
    const float m = b1 * mom[idx] + (1.f - b1) * g;
    const float v = b2 * vel[idx] + (1.f - b2) * g * g;
    mom[idx] = m;
    vel[idx] = v;
    p -= lr * (m / bc1) / (sqrtf(v / bc2) + eps);
    prm[idx] = p;
  }
  if (idx < n1) {
    const uint32_t j = idx / (D + 1), i = idx % (D + 1);
    if (i < D) W1b[size_t(j) * D + i] = f2bf(p);
  } else {
    const uint32_t r = idx - n1;
    const uint32_t k = r / (MLP_HID + 1), j = r % (MLP_HID + 1);
    if (j < MLP_HID) {
      W2b[size_t(k) * MLP_HID + j] = f2bf(p);
      W2Tb[size_t(j) * MLP_CODE + k] = f2bf(p);
    }
  }
}

// Global scale of the sample: partials of sum (x - mu)^2 over 1024-row chunks (FP64, fixed
// order), then s = 1 / sqrt(mean)
__global__ __launch_bounds__(256) void k_scale_part(const float* __restrict__ X, uint32_t N,
                                                    uint32_t D, const float* __restrict__ mu,
                                                    double* part) {
  __shared__ double red[256];
  const uint32_t r0 = blockIdx.x * 1024, r1 = min(N, r0 + 1024);
  double acc = 0.0;
  for (size_t e = size_t(r0) * D + threadIdx.x; e < size_t(r1) * D; e += 256) {
    const double v = double(X[e]) - double(mu[e % D]);
    acc += v * v;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (int(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
__global__ void k_scale_final(const double* part, uint32_t nb, uint32_t N, uint32_t D,
                              float* scal) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double t = 0.0;
  for (uint32_t b = 0; b < nb; ++b) t += part[b];
  const double var = t / (double(N) * D);
  scal[0] = var > 0.0 ? float(1.0 / sqrt(var)) : 1.0f;
}

// batch rows: anchors = samp[perm(step * B + i)], positives = samp[subG[anchor][0..P)]
__global__ void k_mlp_rows(uint32_t* rows, const uint32_t* __restrict__ samp,
                           const uint32_t* __restrict__ subG, uint32_t Ks, uint32_t S,
                           uint32_t B, uint32_t P, uint64_t pa, uint64_t pc, uint32_t step) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B) return;
  const uint32_t j = uint32_t((pa * (uint64_t(step) * B + i) + pc) % S);
  rows[i] = samp[j];
  for (uint32_t p = 0; p < P; ++p) rows[B + i * P + p] = samp[subG[size_t(j) * Ks + p]];
}

// ------------------------------------------------------------------ k-means in code space
__global__ void k_km_init(const float* __restrict__ Z, const uint32_t* __restrict__ samp,
                          uint32_t C, float* cent) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= C * MLP_CODE) return;
  const uint32_t c = e / MLP_CODE, k = e % MLP_CODE;
  cent[e] = Z[size_t(samp[c]) * MLP_CODE + k];
}

__global__ void k_km_norms(const float* __restrict__ cent, uint32_t C, float* cn, uint16_t* chi,
                           uint16_t* clo) {
  const uint32_t c = blockIdx.x * 256 + threadIdx.x;
  if (c >= C) return;
  float s = 0.f;
  for (uint32_t k = 0; k < MLP_CODE; ++k) {
    const float x = cent[c * MLP_CODE + k];
    s = fmaf(x, x, s);
    // the split the matrix-core assignment multiplies: x = hi + lo to about 16 bits
    const uint16_t h = f2bf(x);
    chi[c * MLP_CODE + k] = h;
    clo[c * MLP_CODE + k] = f2bf(x - bf2f(h));
  }
  cn[c] = s;
}

// (s1, c1) <= (s2, c2): the two smallest (s, c) pairs seen, kept with selects (no branches).
// `ordered`: the pairs arrive in increasing c, so an equal s never displaces a kept one.
template <bool ordered>
__device__ __forceinline__ void km_top2(float s, uint32_t c, float& s1, uint32_t& c1, float& s2,
                                        uint32_t& c2) {
  const bool b1 = ordered ? s < s1 : (s < s1 || (s == s1 && c < c1));
  const bool b2 = ordered ? s < s2 : (s < s2 || (s == s2 && c < c2));
  const float ns2 = b1 ? s1 : (b2 ? s : s2);
  const uint32_t nc2 = b1 ? c1 : (b2 ? c : c2);
  s1 = b1 ? s : s1;
  c1 = b1 ? c : c1;
  s2 = ns2;
  c2 = nc2;
}

// Nearest and second-nearest centroid of each code (idx: codes idx[0..n), else 0..n), by
// s = |c|^2 - 2 c.z and then the centroid index.  C[c][j] = cent_c . z_j
// as a 16x16x32 BF16 product per tile (centroids as the A rows, codes as the B columns), each
// FP32 operand split into a BF16 high part and a BF16 remainder, and the three products
// hi.hi + hi.lo + lo.hi summed in FP32 (about 16 mantissa bits of each operand).  A lane
// holds one code column and four centroid rows of a tile, so it keeps its code's two
// nearest centroids by (s, c) with no exchange until the end.  16 QN codes per wave (QN column
// tiles), 64 QN per workgroup; the centroids' hi / lo parts (k_km_norms) are staged in LDS 64
// rows at a time.
template <int QN>
__global__ __launch_bounds__(256) void k_km_assign_mfma(const float* __restrict__ Z,
                                                        const uint32_t* __restrict__ idx,
                                                        uint32_t n, const uint16_t* __restrict__ chi,
                                                        const uint16_t* __restrict__ clo,
                                                        const float* __restrict__ cn, uint32_t C,
                                                        uint32_t* lab, uint32_t* lab2) {
  static_assert(MLP_CODE == 32, "one 16x16x32 K-step per tile");
  constexpr int CP = MLP_CODE + 8;  // padded LDS row (80 bytes)
  __shared__ __attribute__((aligned(16))) uint16_t shi[64 * CP];
  __shared__ __attribute__((aligned(16))) uint16_t slo[64 * CP];
  __shared__ float sn[64];
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, g = l >> 4, t = l & 15;
  // B fragments: code j = base + 16q + t, elements 8g .. 8g+7
  const uint32_t base = blockIdx.x * (64 * QN) + 16 * QN * w;
  bf16x8 bh[QN], bl[QN];
#pragma unroll
  for (int q = 0; q < QN; ++q) {
    const uint32_t j = base + 16 * q + t;
    const uint32_t row = j < n ? (idx ? idx[j] : j) : 0u;
    const float4* zp = reinterpret_cast<const float4*>(Z + size_t(row) * MLP_CODE + 8 * g);
    const float4 z0 = zp[0], z1 = zp[1];
    const float x[8] = {z0.x, z0.y, z0.z, z0.w, z1.x, z1.y, z1.z, z1.w};
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const uint16_t h = f2bf(x[e]);
      bh[q][e] = __builtin_bit_cast(__bf16, h);
      bl[q][e] = __builtin_bit_cast(__bf16, f2bf(x[e] - bf2f(h)));
    }
  }
  float s1[QN], s2[QN];
  uint32_t c1[QN], c2[QN];
#pragma unroll
  for (int q = 0; q < QN; ++q) {
    s1[q] = s2[q] = INFINITY;
    c1[q] = c2[q] = 0xFFFFFFFFu;
  }
  for (uint32_t c0 = 0; c0 < C; c0 += 64) {
    __syncthreads();
    {
      const int r = tid >> 2, pc = tid & 3;  // 64 rows x 4 pieces of 16 bytes
      uint4 vh = make_uint4(0u, 0u, 0u, 0u), vl = make_uint4(0u, 0u, 0u, 0u);
      if (c0 + r < C) {
        vh = *reinterpret_cast<const uint4*>(chi + size_t(c0 + r) * MLP_CODE + 8 * pc);
        vl = *reinterpret_cast<const uint4*>(clo + size_t(c0 + r) * MLP_CODE + 8 * pc);
      }
      *reinterpret_cast<uint4*>(shi + r * CP + 8 * pc) = vh;
      *reinterpret_cast<uint4*>(slo + r * CP + 8 * pc) = vl;
      if (tid < 64) sn[tid] = c0 + tid < C ? cn[c0 + tid] : INFINITY;
    }
    __syncthreads();
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
      const bf16x8 ah = *reinterpret_cast<const bf16x8*>(shi + (16 * ct + t) * CP + 8 * g);
      const bf16x8 al = *reinterpret_cast<const bf16x8*>(slo + (16 * ct + t) * CP + 8 * g);
      float snr[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) snr[r] = sn[16 * ct + 4 * g + r];
#pragma unroll
      for (int q = 0; q < QN; ++q) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[q], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[q], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[q], acc, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 4; ++r)
          km_top2<true>(snr[r] - 2.f * acc[r], c0 + uint32_t(16 * ct + 4 * g + r), s1[q], c1[q],
                        s2[q], c2[q]);
      }
    }
  }
  // the four lane groups hold disjoint centroid rows of each code: merge their top-2
#pragma unroll
  for (int q = 0; q < QN; ++q) {
#pragma unroll
    for (int o = 16; o <= 32; o <<= 1) {
      const float ps1 = __shfl_xor(s1[q], o, TG_WAVE), ps2 = __shfl_xor(s2[q], o, TG_WAVE);
      const uint32_t pc1 = uint32_t(__shfl_xor(int(c1[q]), o, TG_WAVE));
      const uint32_t pc2 = uint32_t(__shfl_xor(int(c2[q]), o, TG_WAVE));
      km_top2<false>(ps1, pc1, s1[q], c1[q], s2[q], c2[q]);
      km_top2<false>(ps2, pc2, s1[q], c1[q], s2[q], c2[q]);
    }
    const uint32_t j = base + 16 * q + t;
    if (g == 0 && j < n) {
      lab[j] = c1[q];
      if (lab2) lab2[j] = C > 1 ? c2[q] : c1[q];
    }
  }
}

__global__ __launch_bounds__(64) void k_km_update(const float* __restrict__ Z,
                                                  const uint32_t* __restrict__ samp,
                                                  const uint32_t* __restrict__ slab, uint32_t S,
                                                  float* cent) {
  const uint32_t c = blockIdx.x;
  const int lane = threadIdx.x;
  float acc[MLP_CODE];
#pragma unroll
  for (int k = 0; k < int(MLP_CODE); ++k) acc[k] = 0.f;
  float cnt = 0.f;
  // each lane visits its members j = lane, lane + 64, ... in ascending order; the labels
  // are loaded eight at a time so that the scan is not one dependent load per step
  for (uint32_t j0 = lane; j0 < S; j0 += 8 * TG_WAVE) {
    uint32_t lb[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const uint32_t j = j0 + uint32_t(r) * TG_WAVE;
      lb[r] = j < S ? slab[j] : 0xFFFFFFFFu;
    }
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      if (lb[r] != c) continue;
      const float* zp = Z + size_t(samp[j0 + uint32_t(r) * TG_WAVE]) * MLP_CODE;
#pragma unroll
      for (int k = 0; k < int(MLP_CODE); ++k) acc[k] += zp[k];
      cnt += 1.f;
    }
  }
  cnt = wave_sum(cnt);
  float mine = 0.f;
#pragma unroll
  for (int k = 0; k < int(MLP_CODE); ++k) {
    const float v = wave_sum(acc[k]);
    if (lane == k) mine = v;
  }
  if (cnt > 0.f && lane < int(MLP_CODE)) cent[size_t(c) * MLP_CODE + lane] = mine / cnt;
}

// ------------------------------------------ stable grouping of the vectors by bucket label
__global__ __launch_bounds__(256) void k_grp_hist(const uint32_t* __restrict__ lab, uint32_t N,
                                                  uint32_t C, uint32_t CH, uint32_t* cnt) {
  __shared__ uint32_t h[MAX_BUCKETS];
  for (uint32_t b = threadIdx.x; b < C; b += 256) h[b] = 0;
  __syncthreads();
  const uint32_t u0 = blockIdx.x * CH, u1 = min(N, u0 + CH);
  for (uint32_t u = u0 + threadIdx.x; u < u1; u += 256) atomicAdd(&h[lab[u]], 1u);
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < C; b += 256) cnt[size_t(blockIdx.x) * C + b] = h[b];
}
// per bucket: exclusive offsets over chunks (in place) and totals
__global__ void k_grp_scan(uint32_t* cnt, uint32_t nch, uint32_t C, uint32_t* tot) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= C) return;
  uint32_t run = 0, ch = 0;
  for (; ch + 16 <= nch; ch += 16) {  // 16 counts in flight per lane
    uint32_t t[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) t[q] = cnt[size_t(ch + q) * C + b];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      cnt[size_t(ch + q) * C + b] = run;
      run += t[q];
    }
  }
  for (; ch < nch; ++ch) {
    const uint32_t t = cnt[size_t(ch) * C + b];
    cnt[size_t(ch) * C + b] = run;
    run += t;
  }
  tot[b] = run;
}
// bucket starts (exclusive scan of the totals, one block); start[C] = N
__global__ __launch_bounds__(1024) void k_grp_start(const uint32_t* __restrict__ tot, uint32_t C,
                                                    uint32_t* start) {
  __shared__ uint32_t part[1024];
  const uint32_t t = threadIdx.x;
  const uint32_t per = (C + 1023) / 1024;
  const uint32_t b0 = min(C, t * per), b1 = min(C, b0 + per);
  uint32_t s = 0;
  for (uint32_t b = b0; b < b1; ++b) s += tot[b];
  part[t] = s;
  __syncthreads();
  if (t == 0) {
    uint32_t run = 0;
    for (int i = 0; i < 1024; ++i) {
      const uint32_t v = part[i];
      part[i] = run;
      run += v;
    }
    start[C] = run;
  }
  __syncthreads();
  uint32_t run = part[t];
  for (uint32_t b = b0; b < b1; ++b) {
    start[b] = run;
    run += tot[b];
  }
}
// order[start[b] + rank] = u with rank = number of smaller ids labelled b (one wave per chunk)
__global__ __launch_bounds__(64) void k_grp_scatter(const uint32_t* __restrict__ lab, uint32_t N,
                                                    uint32_t C, uint32_t CH,
                                                    const uint32_t* __restrict__ off,
                                                    const uint32_t* __restrict__ start,
                                                    uint32_t* order) {
  __shared__ uint32_t base[MAX_BUCKETS];
  const int lane = threadIdx.x;
  for (uint32_t b = lane; b < C; b += TG_WAVE)
    base[b] = start[b] + off[size_t(blockIdx.x) * C + b];
  __syncthreads();
  const uint32_t u0 = blockIdx.x * CH, u1 = min(N, u0 + CH);
  const uint64_t lt = (1ull << lane) - 1ull;
  for (uint32_t s0 = u0; s0 < u1; s0 += TG_WAVE) {
    const uint32_t u = s0 + lane;
    const bool valid = u < u1;
    const uint32_t b = valid ? lab[u] : 0xFFFFFFFFu;
    uint64_t todo = ballot(valid);
    while (todo) {
      const int leader = __ffsll((unsigned long long)todo) - 1;
      const uint32_t lb = __shfl(b, leader, TG_WAVE);
      const uint64_t m = ballot(valid && b == lb);
      const uint32_t bb = base[lb];
      if (valid && b == lb) order[bb + __popcll(m & lt)] = u;
      __syncthreads();
      if (lane == leader) base[lb] = bb + uint32_t(__popcll(m));
      __syncthreads();
      todo &= ~m;
    }
  }
}

// -------------------------------------------------------------------------- list seeding
// One wave per node: K/2 random members of bucket lab[u], then members of bucket lab2[u],
// then random global ids; all distinct and != u.  Draws are accepted in stream order like
// the random init; a bounded number of rounds falls back to a deterministic scan.
__global__ __launch_bounds__(64) void k_seed_lists(const uint32_t* __restrict__ lab,
                                                   const uint32_t* __restrict__ lab2,
                                                   const uint32_t* __restrict__ start,
                                                   const uint32_t* __restrict__ order, uint32_t N,
                                                   uint32_t K, uint64_t seed, uint32_t* G) {
  __shared__ uint32_t acc[128];
  __shared__ uint32_t first[TG_WAVE];
  const uint32_t u = blockIdx.x;
  if (u >= N) return;
  const int lane = threadIdx.x;
  int cnt = 0;
  auto chosen = [&](uint32_t c) {
    bool hit = false;
    for (int j = 0; j < cnt; ++j) hit |= acc[j] == c;
    return hit;
  };
  // members mem[0..n) (mem == null: the ids 0..n-1); `has_u`: u is among them
  auto take = [&](const uint32_t* mem, uint32_t n, bool has_u, int need, uint64_t st0) {
    if (need <= 0 || n == 0) return;
    const uint32_t avail = n - (has_u ? 1u : 0u);
    const int target = cnt + min(need, int(min(avail, 128u)));
    int round = 0;
    if (uint32_t(need) < avail) {
      for (uint64_t base = 0; cnt < target && round < 64; base += TG_WAVE, ++round) {
        uint64_t s = st0 + (base + lane) * 0x9E3779B97F4A7C15ull;
        const uint32_t ix = sm_bounded(sm_next(s), n);
        const uint32_t c = mem ? mem[ix] : ix;
        // a draw repeating one in a lower lane is rejected: sort (id, lane) and keep the
        // first of each id
        uint64_t kv[1] = {(uint64_t(c) << 32) | uint32_t(lane)};
        reg_bitonic<1>(kv);
        const uint32_t prev = uint32_t(__shfl(int(kv[0] >> 32), (lane + 63) & 63, TG_WAVE));
        first[uint32_t(kv[0])] = (lane == 0 || prev != uint32_t(kv[0] >> 32)) ? 1u : 0u;
        __syncthreads();
        const bool ok = c != u && first[lane] != 0 && !chosen(c);
        const uint64_t m = ballot(ok);
        const int rk = cnt + mask_rank(m);
        __syncthreads();
        if (ok && rk < target) acc[rk] = c;
        cnt = min(target, cnt + int(__popcll(m)));
        __syncthreads();
      }
    }
    // take-all (or the fallback after the bounded rounds): scan in order
    for (uint32_t base = 0; base < n && cnt < target; base += TG_WAVE) {
      const uint32_t ix = base + lane;
      const uint32_t c = ix < n ? (mem ? mem[ix] : ix) : 0xFFFFFFFFu;
      const bool ok = ix < n && c != u && !chosen(c);
      const uint64_t m = ballot(ok);
      const int rk = cnt + mask_rank(m);
      __syncthreads();
      if (ok && rk < target) acc[rk] = c;
      cnt = min(target, cnt + int(__popcll(m)));
      __syncthreads();
    }
  };
  const uint32_t b1 = lab[u], b2 = lab2[u];
  const uint64_t st = sm_stream(seed ^ 0x5EED11E5ull, u);
  take(order + start[b1], start[b1 + 1] - start[b1], true, int(K / 2), st);
  if (b2 != b1)
    take(order + start[b2], start[b2 + 1] - start[b2], false, int(K) - cnt,
         st + 0x100000ull * 0x9E3779B97F4A7C15ull);
  take(nullptr, N, true, int(K) - cnt, st + 0x200000ull * 0x9E3779B97F4A7C15ull);
  for (int j = lane; j < int(K); j += TG_WAVE) G[size_t(u) * K + j] = acc[j];
}

// loss of the batch: fixed-order sum of the per-anchor partials (one thread)
__global__ void k_loss_sum(const float* lossp, uint32_t B, float* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  for (uint32_t b = 0; b < B; ++b) s += lossp[b];
  *out = float(s);
}

__global__ void k_fill_u16(uint16_t* p, uint16_t v, size_t n) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <class T>
hipError_t rmalloc(T** p, size_t n) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  return hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(n, 1) * sizeof(T));
}
template <class T>
void rfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

#define LX(expr)                       \
  do {                                 \
    hipError_t e_ = (expr);            \
    if (e_ != hipSuccess) return e_;   \
  } while (0)

uint64_t gcd64(uint64_t a, uint64_t b) {
  while (b) {
    const uint64_t t = a % b;
    a = b;
    b = t;
  }
  return a;
}
uint64_t host_splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ void k_sample_ids(uint32_t* samp, uint32_t S, uint32_t N, uint64_t a, uint64_t b) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < S) samp[j] = uint32_t((a * j + b) % N);
}

__global__ void k_gather_rows(const float* __restrict__ X, uint32_t D,
                              const uint32_t* __restrict__ ids, uint32_t n, float* out) {
  const size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= size_t(n) * D) return;
  const size_t r = e / D, d = e % D;
  out[e] = X[size_t(ids[r]) * D + d];
}

}  // namespace

void launch_sample_ids(uint32_t* samp, uint32_t S, uint32_t N, uint64_t seed, hipStream_t st) {
  // the first S terms of j -> (a j + b) mod N with a coprime to N: distinct ids
  uint64_t rs = seed ^ 0x5A3B1E00ull;
  uint64_t a = host_splitmix(rs) % N, b = host_splitmix(rs) % N;
  if (a == 0) a = 1;
  while (gcd64(a, N) != 1) a = a + 1 == N ? 1 : a + 1;
  k_sample_ids<<<(S + 255) / 256, 256, 0, st>>>(samp, S, N, a, b);
}

void launch_gather_rows(const float* X, uint32_t D, const uint32_t* ids, uint32_t n, float* out,
                        hipStream_t st) {
  const size_t tot = size_t(n) * D;
  k_gather_rows<<<uint32_t((tot + 255) / 256), 256, 0, st>>>(X, D, ids, n, out);
}

hipError_t learner_reserve(Learner& L, uint32_t D, uint32_t M, uint32_t N, uint32_t S, uint32_t C) {
  if (L.D != D) {
    const size_t n = size_t(mlp_n1(D)) + mlp_n2();
    LX(rmalloc(&L.prm, n));
    LX(rmalloc(&L.mom, n));
    LX(rmalloc(&L.vel, n));
    LX(rmalloc(&L.gpart, n * L.splits));
    LX(rmalloc(&L.W1b, size_t(MLP_HID) * D));
    LX(rmalloc(&L.W2b, size_t(MLP_CODE) * MLP_HID));
    LX(rmalloc(&L.W2Tb, size_t(MLP_HID) * MLP_CODE));
    LX(rmalloc(&L.mu, D));
    LX(rmalloc(&L.scal, 4));
    L.D = D;
    L.Mcap = 0;
  }
  const uint32_t rows_needed = std::max({M, S, std::min(N, ENC_CHUNK)});
  if (L.dpart_rows < rows_needed || !L.dpart) {
    LX(rmalloc(&L.dpart, size_t((rows_needed + 1023) / 1024) * std::max(D, 1u)));
    L.dpart_rows = rows_needed;
  }
  const uint32_t Mb = std::max(M, std::min(N, ENC_CHUNK));
  if (L.Mcap < Mb) {
    LX(rmalloc(&L.rows, Mb));
    LX(rmalloc(&L.Xn, size_t(Mb) * D));
    LX(rmalloc(&L.XnT, size_t(D + 1) * Mb));
    LX(rmalloc(&L.Hb, size_t(Mb) * MLP_HID));
    LX(rmalloc(&L.HT, size_t(MLP_HID + 1) * Mb));
    LX(rmalloc(&L.dZb, size_t(Mb) * MLP_CODE));
    LX(rmalloc(&L.dZT, size_t(MLP_CODE) * Mb));
    LX(rmalloc(&L.dHT, size_t(MLP_HID) * Mb));
    LX(rmalloc(&L.Z, size_t(Mb) * MLP_CODE));
    LX(rmalloc(&L.lossp, Mb));
    L.Mcap = Mb;
    L.ones_M = 0;
  }
  if (L.Ncap < N) {
    LX(rmalloc(&L.codes, size_t(N) * MLP_CODE));
    LX(rmalloc(&L.lab, N));
    LX(rmalloc(&L.lab2, N));
    LX(rmalloc(&L.order, N));
    L.Ncap = N;
    L.Ccap = 0;
  }
  if (L.Scap < S) {
    LX(rmalloc(&L.slab, S));
    L.Scap = S;
  }
  const uint32_t nch = (N + grp_chunk(N) - 1) / grp_chunk(N);
  if (L.Ccap < C) {
    LX(rmalloc(&L.cent, size_t(C) * MLP_CODE));
    LX(rmalloc(&L.cn, C));
    LX(rmalloc(&L.chi, size_t(C) * MLP_CODE));
    LX(rmalloc(&L.clo, size_t(C) * MLP_CODE));
    LX(rmalloc(&L.gcnt, size_t(nch) * C));
    LX(rmalloc(&L.gtot, C));
    LX(rmalloc(&L.gstart, size_t(C) + 1));
    L.Ccap = C;
  }
  return hipSuccess;
}

void learner_release(Learner& L) {
  if (L.train_exec) (void)hipGraphExecDestroy(L.train_exec);
  rfree(L.loss_hist);
  rfree(L.prm);
  rfree(L.mom);
  rfree(L.vel);
  rfree(L.gpart);
  rfree(L.W1b);
  rfree(L.W2b);
  rfree(L.W2Tb);
  rfree(L.mu);
  rfree(L.scal);
  rfree(L.dpart);
  rfree(L.rows);
  rfree(L.Xn);
  rfree(L.XnT);
  rfree(L.Hb);
  rfree(L.HT);
  rfree(L.dZb);
  rfree(L.dZT);
  rfree(L.dHT);
  rfree(L.Z);
  rfree(L.lossp);
  rfree(L.codes);
  rfree(L.cent);
  rfree(L.cn);
  rfree(L.chi);
  rfree(L.clo);
  rfree(L.slab);
  rfree(L.lab);
  rfree(L.lab2);
  rfree(L.gcnt);
  rfree(L.gtot);
  rfree(L.gstart);
  rfree(L.order);
  L = Learner{};
}

void launch_mlp_stats(Learner& L, const float* Xs, uint32_t S, uint32_t D, hipStream_t st) {
  launch_centroid(Xs, S, D, L.dpart, L.mu, st);
  const uint32_t nb = (S + 1023) / 1024;
  k_scale_part<<<nb, 256, 0, st>>>(Xs, S, D, L.mu, L.dpart);
  k_scale_final<<<1, 64, 0, st>>>(L.dpart, nb, S, D, L.scal);
}

hipError_t mlp_set_params(Learner& L, const float* host_prm, hipStream_t st) {
  const size_t n = size_t(mlp_n1(L.D)) + mlp_n2();
  LX(hipMemcpyAsync(L.prm, host_prm, n * 4, hipMemcpyHostToDevice, st));
  fill_u32(reinterpret_cast<uint32_t*>(L.mom), 0, n, st);
  fill_u32(reinterpret_cast<uint32_t*>(L.vel), 0, n, st);
  LearnConfig c;
  launch_mlp_adam(L, 0, 0, c, nullptr, st);
  return hipSuccess;
}

void launch_mlp_adam(Learner& L, int mode, uint32_t t, const LearnConfig& c, float* grad_out,
                     hipStream_t st) {
  const uint32_t n = mlp_n1(L.D) + mlp_n2();
  const float bc1 = t ? float(1.0 - std::pow(double(c.beta1), double(t))) : 1.f;
  const float bc2 = t ? float(1.0 - std::pow(double(c.beta2), double(t))) : 1.f;
  k_mlp_adam<<<(n + 255) / 256, 256, 0, st>>>(L.prm, L.mom, L.vel, L.gpart, L.D, L.splits, mode,
                                              c.lr, c.beta1, c.beta2, c.eps, bc1, bc2, L.W1b,
                                              L.W2b, L.W2Tb, grad_out);
}

void mlp_forward_backward(Learner& L, const float* X, const uint32_t* rows, uint32_t B,
                          uint32_t P, float margin, hipStream_t st) {
  const uint32_t D = L.D, M = B * (1 + P);
  // row of ones closing Xn^T and H^T (bias gradients); unchanged while M is
  if (L.ones_M != M) {
    k_fill_u16<<<(M + 255) / 256, 256, 0, st>>>(L.XnT + size_t(D) * M, 0x3F80, M);
    k_fill_u16<<<(M + 255) / 256, 256, 0, st>>>(L.HT + size_t(MLP_HID) * M, 0x3F80, M);
    L.ones_M = M;
  }
  k_mlp_gather<<<dim3((M + 63) / 64, (D + 63) / 64), 256, 0, st>>>(X, D, rows, 0, M, L.mu, L.scal,
                                                                   L.Xn, L.XnT, M);
  GemmArgs g{};
  // H = relu(Xn W1^T + b1)
  g.A = L.Xn; g.lda = D; g.B = L.W1b; g.ldb = D; g.M = M; g.Nc = MLP_HID; g.Kd = D;
  g.Cb = L.Hb; g.ldcb = MLP_HID; g.CT = L.HT; g.ldct = M; g.bias = L.prm + D; g.bias_stride = D + 1;
  gemm<EPI_H>(g, 1, st);
  // Z = H W2^T + b2
  g = GemmArgs{};
  g.A = L.Hb; g.lda = MLP_HID; g.B = L.W2b; g.ldb = MLP_HID; g.M = M; g.Nc = MLP_CODE;
  g.Kd = MLP_HID; g.Cf = L.Z; g.ldc = MLP_CODE; g.bias = L.prm + mlp_n1(D) + MLP_HID;
  g.bias_stride = MLP_HID + 1;
  gemm<EPI_Z>(g, 1, st);
  k_mlp_loss<<<(B + LOSS_AP - 1) / LOSS_AP, 256, 0, st>>>(L.Z, B, P, margin, M, L.dZb, L.dZT,
                                                         L.lossp);
  const uint32_t kt = M / GK;
  const uint32_t splits = std::min(L.splits, std::max(1u, kt));
  // unused split slots must read as zero in the reduction
  if (splits < L.splits)
    fill_u32(reinterpret_cast<uint32_t*>(L.gpart), 0,
             size_t(L.splits) * (size_t(mlp_n1(D)) + mlp_n2()), st);
  // [dW2 | db2] = dZ^T [H | 1]
  g = GemmArgs{};
  g.A = L.dZT; g.lda = M; g.B = L.HT; g.ldb = M; g.M = MLP_CODE; g.Nc = MLP_HID + 1; g.Kd = M;
  g.Cf = L.gpart + size_t(L.splits) * mlp_n1(D); g.ldc = MLP_HID + 1; g.zstride = mlp_n2();
  gemm<EPI_PART>(g, splits, st);
  // dH = (dZ W2) * relu'(H), stored transposed
  g = GemmArgs{};
  g.A = L.dZb; g.lda = MLP_CODE; g.B = L.W2Tb; g.ldb = MLP_CODE; g.M = M; g.Nc = MLP_HID;
  g.Kd = MLP_CODE; g.CT = L.dHT; g.ldct = M; g.mask = L.Hb; g.ldm = MLP_HID;
  gemm<EPI_DH>(g, 1, st);
  // [dW1 | db1] = dH^T [Xn | 1]
  g = GemmArgs{};
  g.A = L.dHT; g.lda = M; g.B = L.XnT; g.ldb = M; g.M = MLP_HID; g.Nc = D + 1; g.Kd = M;
  g.Cf = L.gpart; g.ldc = D + 1; g.zstride = mlp_n1(D);
  gemm<EPI_PART>(g, splits, st);
}

void mlp_encode(Learner& L, const float* X, uint32_t first, uint32_t n, float* out,
                hipStream_t st) {
  const uint32_t D = L.D;
  for (uint32_t c0 = 0; c0 < n; c0 += L.Mcap) {
    const uint32_t m = std::min(L.Mcap, n - c0);
    k_mlp_gather<<<dim3((m + 63) / 64, (D + 63) / 64), 256, 0, st>>>(
        X, D, nullptr, first + c0, m, L.mu, L.scal, L.Xn, nullptr, 0);
    GemmArgs g{};
    g.A = L.Xn; g.lda = D; g.B = L.W1b; g.ldb = D; g.M = m; g.Nc = MLP_HID; g.Kd = D;
    g.Cb = L.Hb; g.ldcb = MLP_HID; g.CT = nullptr; g.bias = L.prm + D; g.bias_stride = D + 1;
    gemm<EPI_H>(g, 1, st);
    g = GemmArgs{};
    g.A = L.Hb; g.lda = MLP_HID; g.B = L.W2b; g.ldb = MLP_HID; g.M = m; g.Nc = MLP_CODE;
    g.Kd = MLP_HID; g.Cf = out + size_t(c0) * MLP_CODE; g.ldc = MLP_CODE;
    g.bias = L.prm + mlp_n1(D) + MLP_HID; g.bias_stride = MLP_HID + 1;
    gemm<EPI_Z>(g, 1, st);
  }
}

hipError_t learned_seed(Learner& L, const float* X, uint32_t N, uint32_t D, uint32_t K,
                        const float* Xs, const uint32_t* samp, uint32_t S, const uint32_t* subG,
                        uint32_t Ks,
                        const LearnConfig& c, uint64_t seed, uint32_t* G, hipStream_t st,
                        LearnStats* stats, bool timed, bool use_graph) {
  using clk = std::chrono::steady_clock;
  auto now_ms = [&](clk::time_point t0) -> double {
    if (!timed) return 0.0;
    (void)hipStreamSynchronize(st);
    return std::chrono::duration<double, std::milli>(clk::now() - t0).count();
  };
  uint32_t B = std::min(c.B, S) / 32 * 32;
  if (B < 32) B = 32;  // callers guarantee S >= 32
  const uint32_t P = std::min(c.P, Ks);
  const uint32_t M = B * (1 + P);
  const uint32_t C = std::min(std::max(1u, N / std::max(1u, c.bucket)), std::min(MAX_BUCKETS, S));
  LX(learner_reserve(L, D, M, N, S, C));
  // sample statistics and deterministic initial weights (uniform +-1/sqrt(fan-in))
  launch_mlp_stats(L, Xs, S, D, st);
  const size_t n1 = mlp_n1(D), n2 = mlp_n2();
  std::vector<float> h(n1 + n2, 0.f);
  uint64_t rs = seed ^ 0x3A1E5EEDull;
  const float a1 = 1.f / std::sqrt(float(D)), a2 = 1.f / std::sqrt(float(MLP_HID));
  for (uint32_t j = 0; j < MLP_HID; ++j)
    for (uint32_t i = 0; i < D; ++i)
      h[size_t(j) * (D + 1) + i] = a1 * (2.f * float(host_splitmix(rs) >> 40) / 16777216.f - 1.f);
  for (uint32_t k = 0; k < MLP_CODE; ++k)
    for (uint32_t j = 0; j < MLP_HID; ++j)
      h[n1 + size_t(k) * (MLP_HID + 1) + j] =
          a2 * (2.f * float(host_splitmix(rs) >> 40) / 16777216.f - 1.f);
  LX(mlp_set_params(L, h.data(), st));
  LX(hipStreamSynchronize(st));  // h must outlive the asynchronous upload
  auto tt = clk::now();
  // training: every epoch visits the sample in the order of a seeded affine permutation
  const uint32_t steps = S / B;
  std::vector<uint64_t> perm_a(c.epochs), perm_c(c.epochs);
  for (uint32_t ep = 0; ep < c.epochs; ++ep) {
    uint64_t a = (host_splitmix(rs) % S) | 1, pc = host_splitmix(rs) % S;
    while (gcd64(a, S) != 1) a = (a + 2) % S == 0 ? 1 : (a + 2) % S;
    perm_a[ep] = a;
    perm_c[ep] = pc;
  }
  if (!L.loss_hist) LX(rmalloc(&L.loss_hist, 2));
  auto run_training = [&]() {
    uint32_t t = 0;
    for (uint32_t ep = 0; ep < c.epochs; ++ep) {
      for (uint32_t sidx = 0; sidx < steps; ++sidx) {
        k_mlp_rows<<<(B + 255) / 256, 256, 0, st>>>(L.rows, samp, subG, Ks, S, B, P,
                                                   perm_a[ep], perm_c[ep], sidx);
        mlp_forward_backward(L, X, L.rows, B, P, c.margin, st);
        if (sidx + 1 == steps && (ep == 0 || ep + 1 == c.epochs))
          k_loss_sum<<<1, 64, 0, st>>>(L.lossp, B, L.loss_hist + (ep == 0 ? 0 : 1));
        launch_mlp_adam(L, 2, ++t, c, nullptr, st);
      }
    }
    return t;
  };
  uint32_t nsteps = steps * c.epochs;
  if (use_graph) {
    // one capture per (buffers, shapes, seed); later builds replay it
    uint64_t key = 0x6A09E667F3BCC909ull;
    auto mix = [&](uint64_t v) { key = (key ^ v) * 0x100000001B3ull; };
    mix(uint64_t(reinterpret_cast<uintptr_t>(X)));
    mix(uint64_t(reinterpret_cast<uintptr_t>(samp)));
    mix(uint64_t(reinterpret_cast<uintptr_t>(subG)));
    mix(uint64_t(reinterpret_cast<uintptr_t>(L.rows)));
    mix(uint64_t(reinterpret_cast<uintptr_t>(L.prm)));
    mix(uint64_t(reinterpret_cast<uintptr_t>(L.gpart)));
    mix(N); mix(D); mix(S); mix(Ks); mix(B); mix(P); mix(c.epochs); mix(seed);
    uint32_t lrb, mb;
    std::memcpy(&lrb, &c.lr, 4);
    std::memcpy(&mb, &c.margin, 4);
    mix(lrb); mix(mb);
    if (!L.train_exec || L.train_key != key) {
      if (L.train_exec) (void)hipGraphExecDestroy(L.train_exec);
      L.train_exec = nullptr;
      hipGraph_t g = nullptr;
      LX(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      run_training();
      LX(hipStreamEndCapture(st, &g));
      const hipError_t e = hipGraphInstantiate(&L.train_exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      LX(e);
      L.train_key = key;
    }
    LX(hipGraphLaunch(L.train_exec, st));
  } else {
    nsteps = run_training();
  }
  if (stats) {
    float lh[2] = {0.f, 0.f};
    LX(hipMemcpyAsync(lh, L.loss_hist, sizeof(lh), hipMemcpyDeviceToHost, st));
    LX(hipStreamSynchronize(st));
    stats->loss_first = lh[0];
    stats->loss_last = lh[1];
    stats->ms_train = now_ms(tt);
    stats->steps = nsteps;
  }
  auto te = clk::now();
  mlp_encode(L, X, 0, N, L.codes, st);
  if (stats) stats->ms_encode = now_ms(te);
  auto tc = clk::now();
  // k-means on the sample's codes, then the two nearest buckets of every vector
  k_km_init<<<(C * MLP_CODE + 255) / 256, 256, 0, st>>>(L.codes, samp, C, L.cent);
  for (uint32_t it = 0; it < c.kmeans_iters; ++it) {
    k_km_norms<<<(C + 255) / 256, 256, 0, st>>>(L.cent, C, L.cn, L.chi, L.clo);
    k_km_assign_mfma<1><<<(S + 63) / 64, 256, 0, st>>>(L.codes, samp, S, L.chi, L.clo, L.cn, C,
                                                        L.slab, nullptr);
    k_km_update<<<C, 64, 0, st>>>(L.codes, samp, L.slab, S, L.cent);
  }
  k_km_norms<<<(C + 255) / 256, 256, 0, st>>>(L.cent, C, L.cn, L.chi, L.clo);
  k_km_assign_mfma<4><<<(N + 255) / 256, 256, 0, st>>>(L.codes, nullptr, N, L.chi, L.clo, L.cn,
                                                       C, L.lab, L.lab2);
  const uint32_t ch = grp_chunk(N), nch = (N + ch - 1) / ch;
  k_grp_hist<<<nch, 256, 0, st>>>(L.lab, N, C, ch, L.gcnt);
  k_grp_scan<<<(C + 255) / 256, 256, 0, st>>>(L.gcnt, nch, C, L.gtot);
  k_grp_start<<<1, 1024, 0, st>>>(L.gtot, C, L.gstart);
  k_grp_scatter<<<nch, 64, 0, st>>>(L.lab, N, C, ch, L.gcnt, L.gstart, L.order);
  if (stats) {
    stats->ms_cluster = now_ms(tc);
    stats->buckets = C;
  }
  auto ts = clk::now();
  k_seed_lists<<<N, 64, 0, st>>>(L.lab, L.lab2, L.gstart, L.order, N, K, seed, G);
  if (stats) stats->ms_seed = now_ms(ts);
  return hipGetLastError();
}

}  // namespace tg
