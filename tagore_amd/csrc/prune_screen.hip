// prune_screen.hip -- DPG occlusion counts from an MFMA Gram of each node's neighbours.
//
// DPG (filter_dpg, Tagore_src.cu:1713-1785; rule A.9) counts, for every neighbour j of u,
// the earlier neighbours r < j with d(G[u][r], G[u][j]) < D[u][j].  That is K(K-1)/2
// canonical distances per node (2,016 at K=64), which at dimension 960 makes the plain kernel
// read each pair's two rows: ~15 MB per node.  Here the K rows are staged through LDS once,
// their Gram matrix runs on MFMA (BF16 16x16x32, or FP8 e4m3 16x16x32 with per-vector scales),
// and each screen distance s comes with the rigorous bound E of the local join's screen:
//   s + E <  D[u][j]  -> occludes (d <= s + E);
//   s - E >= D[u][j]  -> does not (d >= s - E);
//   otherwise the canonical FP32 distance decides.
// The counts, and so the pruned rows, are those of the exact rule.  K <= 64.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace tg {

namespace {

__device__ __forceinline__ float dpg_bound(float na, float qa, float ea, float nb, float qb,
                                           float eb, float eta) {
  return 2.0f * (ea * qb + qa * eb + ea * eb) + eta * (na + nb + 2.0f * qa * qb);
}

// PREC 2: BF16 rows (a.Xb), PREC 3: FP8 rows (a.Xq8, scales a.sc8).  256 threads; wave w
// owns rows 16w..16w+15 of the 64 x 64 Gram, against the 4 column blocks.
template <int PREC, int MC, class T>
__global__ __launch_bounds__(256) void k_dpg_screen(PruneArgs a, DpgScreen q) {
  constexpr int DC = PREC == 3 ? 256 : 128;  // elements staged per round (256 bytes per row)
  constexpr int ESZ = PREC == 3 ? 1 : 2;
  constexpr int CPR = DC * ESZ / 16;
  constexpr int RS = DC * ESZ + 16;
  constexpr int MROW = 65;
  constexpr int SMEM = 64 * RS > 64 * MROW * 4 ? 64 * RS : 64 * MROW * 4;
  __shared__ __attribute__((aligned(16))) unsigned char smem[SMEM];
  __shared__ uint32_t nei[64], det[64];
  __shared__ float nn[64], nq[64], ne[64], sc[64], thr[64];
  __shared__ uint16_t queue[64 * 63 / 2];
  __shared__ uint32_t nqueue;
  __shared__ uint64_t srt[64];
  float* M = reinterpret_cast<float*>(smem);

  const uint32_t u = row_of_block(a.perm, a.N);
  if (u >= a.N) return;
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  const int K = a.K, D = a.D, d = a.FD;
  if (tid < 64) {
    det[tid] = 0;
    if (tid < K) {
      const uint32_t v = a.G[size_t(u) * K + tid];
      nei[tid] = v;
      nn[tid] = q.nrm2[v];
      nq[tid] = q.nq[v];
      ne[tid] = q.ne[v];
      sc[tid] = PREC == 3 ? q.sc8[v] : 1.f;
      thr[tid] = a.GD[size_t(u) * K + tid];
    }
  }
  if (tid == 0) nqueue = 0;
  __syncthreads();
  f32x4 acc[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) acc[c] = f32x4{0.f, 0.f, 0.f, 0.f};
  const unsigned char* src = PREC == 3 ? reinterpret_cast<const unsigned char*>(q.Xq8)
                                       : reinterpret_cast<const unsigned char*>(q.Xb);
  for (int d0 = 0; d0 < D; d0 += DC) {
    for (int e = tid; e < 64 * CPR; e += 256) {
      const int r = e / CPR, c16 = e % CPR;
      uint4 v = make_uint4(0, 0, 0, 0);
      const int col = d0 + c16 * (16 / ESZ);
      if (r < K && col < D)
        v = *reinterpret_cast<const uint4*>(src + (size_t(nei[r]) * D + col) * ESZ);
      *reinterpret_cast<uint4*>(smem + r * RS + c16 * 16) = v;
    }
    __syncthreads();
#pragma unroll
    for (int k0 = 0; k0 < DC; k0 += 32) {
      const int kk = k0 + 8 * (l >> 4);
      if (PREC == 3) {
        const long fa = *reinterpret_cast<const long*>(smem + (16 * w + (l & 15)) * RS + kk);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const long fc = *reinterpret_cast<const long*>(smem + (16 * c + (l & 15)) * RS + kk);
          acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(fa, fc, acc[c], 0, 0, 0);
        }
      } else {
        const bf16x8 fa = *reinterpret_cast<const bf16x8*>(smem + (16 * w + (l & 15)) * RS + kk * 2);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const bf16x8 fc = *reinterpret_cast<const bf16x8*>(smem + (16 * c + (l & 15)) * RS + kk * 2);
          acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa, fc, acc[c], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
  // screen distances (the matrix aliases the staging area, free after the last round)
#pragma unroll
  for (int c = 0; c < 4; ++c) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = 16 * w + 4 * (l >> 4) + r, j = 16 * c + (l & 15);
      float s = 0.f;
      if (i < K && j < K) s = (nn[i] + nn[j]) - 2.0f * ((sc[i] * sc[j]) * acc[c][r]);
      M[i * MROW + j] = s;
    }
  }
  __syncthreads();
  // classify every pair r < j
  const int npairs = K * (K - 1) / 2;
  for (int p = tid; p < npairs; p += 256) {
    // p -> (r, j), j = 1..K-1, pairs of column j are r = 0..j-1
    int j = int((1.0f + sqrtf(1.0f + 8.0f * float(p))) * 0.5f);
    while (j * (j - 1) / 2 > p) --j;
    while ((j + 1) * j / 2 <= p) ++j;
    const int r = p - j * (j - 1) / 2;
    const float s = M[r * MROW + j];
    const float E = dpg_bound(nn[r], nq[r], ne[r], nn[j], nq[j], ne[j], q.eta);
    const float tau = thr[j];
    if (s + E < tau) {
      atomicAdd(&det[j], 1u);
    } else if (!(s - E >= tau)) {
      const uint32_t slot = atomicAdd(&nqueue, 1u);
      queue[slot] = uint16_t((r << 8) | j);
    }
  }
  __syncthreads();
  // canonical distances of the undecided pairs, four per 16-lane group in flight
  const int nqd = int(nqueue);
  const int gg = tid >> 4, t = tid & 15;
  const T* rows = reinterpret_cast<const T*>(a.rows);
  for (int e0 = 0; e0 < nqd; e0 += 64) {
    const T* ra[4];
    const T* rb[4];
    int jj[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int e = e0 + gg + 16 * b;
      ra[b] = rb[b] = rows;  // unused slots read row 0
      jj[b] = -1;
      if (e < nqd) {
        const int r = queue[e] >> 8, j = queue[e] & 255;
        ra[b] = rows + size_t(nei[r]) * D;
        rb[b] = rows + size_t(nei[j]) * D;
        jj[b] = j;
      }
    }
    float dd[4];
    dist16_pairs<MC, 4>(ra, rb, D, t, dd);
    if (t == 0) {
#pragma unroll
      for (int b = 0; b < 4; ++b)
        if (jj[b] >= 0 && dd[b] < thr[jj[b]]) atomicAdd(&det[jj[b]], 1u);
    }
  }
  __syncthreads();
  if (tid == 0 && a.counters) {
    atomicAdd(&cnt_slot(a.counters)[1], (unsigned long long)npairs);
    atomicAdd(&cnt_slot(a.counters)[5], (unsigned long long)nqd);
  }
  // select the d entries of smallest (detour count, id) and push the reverse pairs
  if (tid < 64) srt[tid] = tid < K ? ((uint64_t(det[tid]) << 32) | nei[tid]) : TG_KEY_INF;
  __syncthreads();
  wave_bitonic(srt, 64);  // every thread calls it: 32 compare-exchanges per step, one each
  for (int i = tid; i < d; i += 256) {
    const uint64_t k = srt[i];
    a.fwd_ids[size_t(u) * d + i] = uint32_t(k);
    pb_push(a.rev, uint32_t(k), (k & 0xFFFFFFFF00000000ull) | u);
  }
  if (tid == 0) a.fwd_deg[u] = d;
}

}  // namespace

void launch_dpg_screen(const PruneArgs& a, const DpgScreen& q, int precision, hipStream_t s) {
  const int mc = (int(a.D) + 63) / 64;
#define DS(P, M, T) k_dpg_screen<P, M, T><<<grid_rows(a.perm, a.N), 256, 0, s>>>(a, q)
#define DS_MC(P, T)                                                    \
  do {                                                                 \
    if (mc <= 2) DS(P, 2, T); else if (mc <= 4) DS(P, 4, T); else DS(P, 16, T); \
  } while (0)
  if (precision == KP_FP8) {
    if (a.rows_bf16) DS_MC(3, uint16_t); else DS_MC(3, float);
  } else {
    if (a.rows_bf16) DS_MC(2, uint16_t); else DS_MC(2, float);
  }
#undef DS_MC
#undef DS
}

}  // namespace tg
