// learn.hpp -- learned initial kNN lists: device buffers and host entry points.
//
// The reference starts GNN-Descent from random lists (initialize_graph, Tagore_src.cu:158-166).
// This module replaces that start, when asked, by lists seeded from an MLP encoder's codes:
//   1. encoder  z = W2 relu(W1 xn + b1) + b2, xn = (x - mu) * s  (D -> 256 -> 32), FP32 master
//      weights with BF16 copies; every product runs on v_mfma_f32_16x16x32_bf16 with FP32
//      accumulation;
//   2. training: Adam on a contrastive loss over a data sample's kNN graph (anchors b, their
//      P nearest sample neighbours p(b,j), all other anchors of the batch as negatives):
//        L = 1/(B P) sum_{b,j} |z_b - z_p(b,j)|^2 + 1/(B (B-1)) sum_{b != c} relu(m - |z_b - z_c|^2)
//   3. k-means in code space (Lloyd on the sample's codes), every vector's two nearest buckets;
//   4. each initial list: K/2 random members of the nearest bucket, the rest from the second
//      nearest (random global ids when the buckets hold too few), canonical FP32 distances.
// Every reduction runs in a fixed order, so the whole path is deterministic.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tg {

constexpr uint32_t MLP_HID = 256;
constexpr uint32_t MLP_CODE = 32;

// flat parameter block: W1|b1 as [HID][D+1] (column D = b1), then W2|b2 as [CODE][HID+1]
__host__ __device__ inline uint32_t mlp_n1(uint32_t D) { return MLP_HID * (D + 1); }
__host__ __device__ inline uint32_t mlp_n2() { return MLP_CODE * (MLP_HID + 1); }

struct Learner {
  uint32_t D = 0, Mcap = 0, splits = 16;
  // parameters, Adam moments, split-K gradient partials, BF16 weight copies
  float *prm = nullptr, *mom = nullptr, *vel = nullptr, *gpart = nullptr;
  uint16_t *W1b = nullptr, *W2b = nullptr, *W2Tb = nullptr;
  float* mu = nullptr;      // [D]
  float* scal = nullptr;    // [1]: s = 1 / global standard deviation of the sample
  double* dpart = nullptr;  // [ceil(rows/1024) * max(D,1)] column / scale partials
  uint32_t dpart_rows = 0;
  // batch (M rows: B anchors, then B*P positives)
  uint32_t* rows = nullptr;
  uint16_t *Xn = nullptr, *XnT = nullptr, *Hb = nullptr, *HT = nullptr;
  uint16_t *dZb = nullptr, *dZT = nullptr, *dHT = nullptr;
  float *Z = nullptr, *lossp = nullptr;
  uint32_t ones_M = 0;  // batch size whose rows of ones are in place in XnT / HT
  // codes and buckets
  float* codes = nullptr;  // [N * CODE]
  uint32_t Ncap = 0, Ccap = 0;
  float *cent = nullptr, *cn = nullptr;
  uint16_t *chi = nullptr, *clo = nullptr;  // centroids as BF16 high parts and remainders
  uint32_t *slab = nullptr, *lab = nullptr, *lab2 = nullptr, *gcnt = nullptr, *gtot = nullptr,
           *gstart = nullptr, *order = nullptr;
  uint32_t Scap = 0;
  // the training loop captured as a HIP graph, replayed while its key (shapes, buffers,
  // seed) is unchanged; loss sums at the end of the first and last epochs
  hipGraphExec_t train_exec = nullptr;
  uint64_t train_key = 0;
  float* loss_hist = nullptr;  // [2]
};

struct LearnConfig {
  uint32_t B = 2048;      // anchors per batch
  uint32_t P = 4;         // positives per anchor
  uint32_t epochs = 2;
  uint32_t kmeans_iters = 5;
  uint32_t bucket = 1024; // target vectors per bucket: C = clamp(N / bucket, 1, 4096)
  float lr = 3e-3f, margin = 1.0f;
  float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
};

struct LearnStats {
  double ms_train = 0, ms_encode = 0, ms_cluster = 0, ms_seed = 0;
  float loss_first = 0, loss_last = 0;
  uint32_t buckets = 0, steps = 0;
};

// Sample ids: the first S terms of a seeded affine permutation of [0, N); rows X[ids] -> out.
void launch_sample_ids(uint32_t* samp, uint32_t S, uint32_t N, uint64_t seed, hipStream_t st);
void launch_gather_rows(const float* X, uint32_t D, const uint32_t* ids, uint32_t n, float* out,
                        hipStream_t st);

// Allocation (idempotent for equal or smaller sizes); release frees everything.
hipError_t learner_reserve(Learner& L, uint32_t D, uint32_t M, uint32_t N, uint32_t S, uint32_t C);
void learner_release(Learner& L);

// Normalisation statistics (mu, s) of `rows` contiguous vectors.
void launch_mlp_stats(Learner& L, const float* Xs, uint32_t S, uint32_t D, hipStream_t st);
// Initialise the parameters from host values (flat block) and clear the Adam moments.
hipError_t mlp_set_params(Learner& L, const float* host_prm, hipStream_t st);

// One forward/backward over a batch whose rows are X[rows[i]] (rows == null: X[i]); rows
// [0,B) are anchors, rows B + b*P + j the positives of anchor b.  Leaves the split-K gradient
// partials in L.gpart and per-anchor losses in L.lossp.
void mlp_forward_backward(Learner& L, const float* X, const uint32_t* rows, uint32_t B,
                          uint32_t P, float margin, hipStream_t st);
// mode 0: refresh the BF16 weight copies; 1: reduce the gradient into grad_out only;
// 2: Adam step (t >= 1) then refresh the copies.
void launch_mlp_adam(Learner& L, int mode, uint32_t t, const LearnConfig& c, float* grad_out,
                     hipStream_t st);
// Codes of X[first .. first+n) into out[n * CODE].
void mlp_encode(Learner& L, const float* X, uint32_t first, uint32_t n, float* out,
                hipStream_t st);

// Train on a sample (samp[S] global ids; subG[S*Ks] its kNN graph in sample indices), encode
// all N vectors, bucket them, and write K seeded ids per row into G (unsorted, no distances).
// Xs: the sample's rows, contiguous [S*D] (normalisation statistics).
hipError_t learned_seed(Learner& L, const float* X, uint32_t N, uint32_t D, uint32_t K,
                        const float* Xs, const uint32_t* samp, uint32_t S, const uint32_t* subG,
                        uint32_t Ks,
                        const LearnConfig& c, uint64_t seed, uint32_t* G, hipStream_t st,
                        LearnStats* stats, bool timed, bool use_graph = false);

}  // namespace tg
