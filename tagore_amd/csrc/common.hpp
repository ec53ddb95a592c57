// common.hpp -- device primitives shared by the Tagore-on-MI355X kernels.
//
// * (distance, id) keys packed into u64 so that unsigned order == (d, id) order for the
//   non-negative distances used everywhere (+inf sorts after every finite distance).
// * Canonical FP32 squared L2 evaluated by a 16-lane group with float4 loads: lane t owns
//   the partial sums p[4t..4t+3] over i == l (mod 64); they are summed inside the lane,
//   (p[4t] + p[4t+1]) + (p[4t+2] + p[4t+3]), then across the group by DPP row shifts 8,4,2,1.
//   This is the same arithmetic, in the same order, as the CPU oracle (bit-identical).
// * Wave-level (64-lane) bitonic sorts and scans in LDS.
// * PushBuf: per-destination append lists with a bounded slot area and a global overflow
//   area; every reduction sorts, so results do not depend on arrival order.
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>
#include <stdint.h>

#define TG_WAVE 64
#define TG_INV 0xFFFFFFFFu
#define TG_KEY_INF 0xFFFFFFFFFFFFFFFFull

namespace tg {

__device__ __forceinline__ uint64_t mk_key(float d, uint32_t id) {
  return (uint64_t(__float_as_uint(d)) << 32) | id;
}
__device__ __forceinline__ float key_d(uint64_t k) { return __uint_as_float(uint32_t(k >> 32)); }
__device__ __forceinline__ uint32_t key_id(uint64_t k) { return uint32_t(k); }
__device__ __forceinline__ uint64_t key_with_id(uint64_t k, uint32_t id) {
  return (k & 0xFFFFFFFF00000000ull) | id;
}
// finite distance (not +inf / invalid key)
__device__ __forceinline__ bool key_finite(uint64_t k) { return uint32_t(k >> 32) < 0x7F800000u; }

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// wave-wide ballot (64-bit mask)
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// number of set bits of m below this lane
__device__ __forceinline__ int mask_rank(uint64_t m) {
  return __popcll(m & ((1ull << lane_id()) - 1ull));
}

// ---------------------------------------------------------------------------------------
// Evaluation counters: CNT_N counters in each of CNT_SLOTS 128-byte slots; workgroup b adds
// to slot b % CNT_SLOTS, so that a million workgroups do not serialise their atomics on one
// cache line.  The host sums the slots.
// ---------------------------------------------------------------------------------------
constexpr int CNT_N = 16, CNT_SLOTS = 64;
__device__ __forceinline__ unsigned long long* cnt_slot(unsigned long long* c) {
  return c + (blockIdx.x % CNT_SLOTS) * CNT_N;
}

// ---------------------------------------------------------------------------------------
// Canonical distance.  `MC` = number of 64-element chunks held (ceil(D/64) <= MC).
// Row a is preloaded by the group (ra[]), row b is read from memory.  The value is valid in
// lane 0 of the group.  Requires D % 4 == 0 (host checks).
// ---------------------------------------------------------------------------------------
template <int MC>
struct Row16 {
  float4 v[MC];  // FP32 rows: elements m*64 + 4t .. +3; u8 rows: 16 bytes (quad layout below)
  int n2;        // u8 rows only: |x|^2 of the row, exact, in every lane
};

// Four consecutive elements as FP32: rows are stored as FP32, or as BF16 bits when every
// value of the data set is exactly representable in BF16 (then the widened values are the
// FP32 values themselves and every distance below is bit-identical).
// SplitMix64 (Steele, Lea, Flood, OOPSLA 2014): step, per-index stream start, and a
// multiply-shift map of the high 32 bits onto [0, n)
__device__ __forceinline__ uint64_t sm_next(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ uint64_t sm_stream(uint64_t seed, uint64_t idx) {
  uint64_t s = seed ^ (idx * 0xD6E8FEB86659FD93ull);
  return sm_next(s);
}
__device__ __forceinline__ uint32_t sm_bounded(uint64_t x, uint32_t n) {
  return uint32_t(((x >> 32) * uint64_t(n)) >> 32);
}

// Row handled by this workgroup.  With an order (perm: a permutation of the rows that puts
// neighbouring rows together), workgroup b takes position (b % 8) * ceil(N/8) + b / 8 of
// it: workgroups are dealt to the 8 XCDs round-robin, so each XCD walks one contiguous run
// of the order and its L2 keeps the rows its neighbours gather.  Rows past N: 0xFFFFFFFF.
__device__ __forceinline__ uint32_t row_of_block(const uint32_t* perm, uint32_t N) {
  const uint32_t b = blockIdx.x;
  if (!perm) return b;
  const uint32_t chunk = (N + 7) / 8;
  const uint32_t pos = (b & 7) * chunk + (b >> 3);
  return pos < N ? perm[pos] : 0xFFFFFFFFu;
}
__host__ __device__ inline uint32_t grid_rows(const uint32_t* perm, uint32_t N) {
  return perm ? 8 * ((N + 7) / 8) : N;
}

// seed of segment c of a segmented handle (host and device)
__host__ __device__ inline uint64_t segment_seed(uint64_t seed, uint32_t c) {
  return seed ^ ((uint64_t(c) + 1) * 0xA0761D6478BD642Full);
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// round-to-nearest-even FP32 -> BF16 bits (finite inputs)
__device__ __forceinline__ uint16_t f2bf(float f) {
  const uint32_t u = __float_as_uint(f);
  return uint16_t((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float(uint32_t(b) << 16); }

__device__ __forceinline__ float4 load4(const float* p) {
  return *reinterpret_cast<const float4*>(p);
}
__device__ __forceinline__ float4 load4(const uint16_t* p) {
  const uint2 q = *reinterpret_cast<const uint2*>(p);
  return make_float4(__uint_as_float(q.x << 16), __uint_as_float(q.x & 0xFFFF0000u),
                     __uint_as_float(q.y << 16), __uint_as_float(q.y & 0xFFFF0000u));
}

// u8 rows use a quad layout: the four lanes j = lane & 3 of every 4-lane quad share one row,
// lane j holding bytes m*64 + 16j .. +15 of each 64-byte chunk m (zero past D).  A row is
// then two 16-byte loads per lane at D = 128, and a dot product ends with two DPP steps
// inside the quad.  All of it is integer arithmetic, so the lane layout does not change a
// result.
__device__ __forceinline__ uint4 load_u8x16(const uint8_t* x, int m, int D, int j) {
  const int i = m * 64 + 16 * j;
  if ((m + 1) * 64 <= D && (D & 15) == 0)  // uniform: whole chunk, 16-byte aligned rows
    return *reinterpret_cast<const uint4*>(x + i);
  uint4 r = make_uint4(0u, 0u, 0u, 0u);
  if (i < D) r.x = *reinterpret_cast<const uint32_t*>(x + i);
  if (i + 4 < D) r.y = *reinterpret_cast<const uint32_t*>(x + i + 4);
  if (i + 8 < D) r.z = *reinterpret_cast<const uint32_t*>(x + i + 8);
  if (i + 12 < D) r.w = *reinterpret_cast<const uint32_t*>(x + i + 12);
  return r;
}
__device__ __forceinline__ uint4 as_u4(const float4& v) {
  return make_uint4(__float_as_uint(v.x), __float_as_uint(v.y), __float_as_uint(v.z),
                    __float_as_uint(v.w));
}
// acc + the dot product of 16 bytes pairs (v_dot4_u32_u8)
__device__ __forceinline__ uint32_t dot16_u8(const uint4& a, const uint4& b, uint32_t acc) {
  acc = __builtin_amdgcn_udot4(a.x, b.x, acc, false);
  acc = __builtin_amdgcn_udot4(a.y, b.y, acc, false);
  acc = __builtin_amdgcn_udot4(a.z, b.z, acc, false);
  return __builtin_amdgcn_udot4(a.w, b.w, acc, false);
}
// Sum of an integer over the 4 lanes of a quad, in every lane of it.
__device__ __forceinline__ int quad_sum(int q) {
  q += __builtin_amdgcn_mov_dpp(q, 0xB1, 0xF, 0xF, false);  // lane ^ 1
  return q + __builtin_amdgcn_mov_dpp(q, 0x4E, 0xF, 0xF, false);  // lane ^ 2
}

// u8 rows, quad layout: v[m] holds the lane's 16 bytes of chunk m as bits.  The loads are
// issued by row16_u8_load and |x|^2 (n2, in every lane) formed by row16_u8_norm, so that
// other work can run while the loads are in flight.
template <int MC>
__device__ __forceinline__ void row16_u8_load(Row16<MC>& r, const uint8_t* __restrict__ x, int D) {
  const int j = threadIdx.x & 3;
#pragma unroll
  for (int m = 0; m < MC; ++m) {
    const uint4 w = m * 64 < D ? load_u8x16(x, m, D, j) : make_uint4(0u, 0u, 0u, 0u);
    r.v[m] = make_float4(__uint_as_float(w.x), __uint_as_float(w.y), __uint_as_float(w.z),
                         __uint_as_float(w.w));
  }
}
template <int MC>
__device__ __forceinline__ void row16_u8_norm(Row16<MC>& r) {
  uint32_t sq = 0;
#pragma unroll
  for (int m = 0; m < MC; ++m) {
    const uint4 w = make_uint4(__float_as_uint(r.v[m].x), __float_as_uint(r.v[m].y),
                               __float_as_uint(r.v[m].z), __float_as_uint(r.v[m].w));
    sq = dot16_u8(w, w, sq);
  }
  r.n2 = quad_sum(int(sq));
}

template <int MC, class T>
__device__ __forceinline__ void load_row16(Row16<MC>& r, const T* __restrict__ x, int D, int t) {
  if constexpr (std::is_same<T, uint8_t>::value) {
    // u8 rows, quad layout (t is not used: the layout follows lane & 3)
    row16_u8_load<MC>(r, x, D);
    row16_u8_norm<MC>(r);
  } else {
#pragma unroll
    for (int m = 0; m < MC; ++m) {
      const int i = m * 64 + 4 * t;
      if (i < D)
        r.v[m] = load4(x + i);
      else
        r.v[m] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    r.n2 = 0;
  }
}

// lane i of each 16-lane DPP row reads lane i+S of the same row (lanes past the row end read
// 0; only lanes whose partner exists are used)
template <int S>
__device__ __forceinline__ float row_shl(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x100 + S, 0xF, 0xF,
                                                    false));
}

// Reduction of the group's partials (lane t holds p[4t..4t+3]): first inside the lane,
// q[t] = (p[4t] + p[4t+1]) + (p[4t+2] + p[4t+3]), then the tree q[t] += q[t+s] for
// s = 8,4,2,1 as DPP row shifts across the group.  Six levels, like any balanced tree of
// the 64 partials.  The result is valid in lane 0 of the group (t == 0) only.
__device__ __forceinline__ float tree16(float p0, float p1, float p2, float p3) {
  float q = (p0 + p1) + (p2 + p3);
  q = q + row_shl<8>(q);
  q = q + row_shl<4>(q);
  q = q + row_shl<2>(q);
  return q + row_shl<1>(q);
}

// Row chunks held in flight per pair / candidate: at most 4 x 64 elements, so wide rows
// (D up to 1024) stream through the same registers in order (each partial still folds its
// elements in ascending index order).
template <int MC>
struct RowChunk {
  static constexpr int N = MC < 4 ? MC : 4;
};

// Squared differences of one chunk pair folded into the four partials.  Elements past D are
// zero in both operands, so they add fmaf(0, 0, p) = p and leave the partials unchanged.
__device__ __forceinline__ void fold4(const float4& x, const float4& y, float (&p)[4]) {
  const float e0 = x.x - y.x, e1 = x.y - y.y, e2 = x.z - y.z, e3 = x.w - y.w;
  p[0] = __builtin_fmaf(e0, e0, p[0]);
  p[1] = __builtin_fmaf(e1, e1, p[1]);
  p[2] = __builtin_fmaf(e2, e2, p[2]);
  p[3] = __builtin_fmaf(e3, e3, p[3]);
}

// Elements m*64 + 4t .. +3 of row x, zero past D.
template <class T>
__device__ __forceinline__ float4 load_chunk_tail(const T* x, int m, int D, int t) {
  const int i = m * 64 + 4 * t;
  return i < D ? load4(x + i) : make_float4(0.f, 0.f, 0.f, 0.f);
}

// NB distances from row a (registers) to rows b[0..NB).  Every b[k] must point at a
// readable row (callers aim unused slots at any valid row and ignore the result).  Chunks
// of 64 elements that lie wholly inside D are loaded and folded without per-lane masks
// (the test is uniform across the wave); only a partial last chunk masks its loads.  The
// loads of a group of chunks are issued before its arithmetic.  Results valid in lane t == 0.
template <int MC, int NB, class T>
__device__ __forceinline__ void dist16_batch(const Row16<MC>& a, const T* const (&b)[NB], int D,
                                             int t, float (&out)[NB]) {
  constexpr int CH = RowChunk<MC>::N;
  float p[NB][4];
#pragma unroll
  for (int k = 0; k < NB; ++k) p[k][0] = p[k][1] = p[k][2] = p[k][3] = 0.f;
#pragma unroll
  for (int m0 = 0; m0 < MC; m0 += CH) {
    if (m0 * 64 >= D) break;
    float4 y[NB][CH];
    if ((m0 + CH) * 64 <= D) {
#pragma unroll
      for (int k = 0; k < NB; ++k)
#pragma unroll
        for (int c = 0; c < CH; ++c) y[k][c] = load4(b[k] + (m0 + c) * 64 + 4 * t);
#pragma unroll
      for (int k = 0; k < NB; ++k)
#pragma unroll
        for (int c = 0; c < CH; ++c) fold4(a.v[m0 + c], y[k][c], p[k]);
    } else {
#pragma unroll
      for (int k = 0; k < NB; ++k)
#pragma unroll
        for (int c = 0; c < CH; ++c) y[k][c] = load_chunk_tail(b[k], m0 + c, D, t);
#pragma unroll
      for (int k = 0; k < NB; ++k)
#pragma unroll
        for (int c = 0; c < CH; ++c)
          if ((m0 + c) * 64 < D) fold4(a.v[m0 + c], y[k][c], p[k]);
    }
  }
#pragma unroll
  for (int k = 0; k < NB; ++k) out[k] = tree16(p[k][0], p[k][1], p[k][2], p[k][3]);
}

// Distance between two u8 rows held in registers (quad layout), in every lane: the integer
// |a|^2 + |b|^2 - 2 a.b with v_dot4_u32_u8, exact.  The u8 rows are used only when every
// value of the data set is an integer in [0, 255] and every squared distance is below 2^24:
// then every partial sum of the canonical FP32 fold is an exact integer too, so both give
// the same float.
template <int MC>
__device__ __forceinline__ float dist_rows_u8(const Row16<MC>& a, const Row16<MC>& b, int D) {
  uint32_t ab = 0;
#pragma unroll
  for (int m = 0; m < MC; ++m)
    if (m * 64 < D) ab = dot16_u8(as_u4(a.v[m]), as_u4(b.v[m]), ab);
  return float(a.n2 + b.n2 - 2 * quad_sum(int(ab)));
}

// NB distances between the row pairs (a[k], b[k]), both read from memory; every pointer
// must be a readable row, as for dist16_batch.
template <int MC, int NB, class T>
__device__ __forceinline__ void dist16_pairs(const T* const (&a)[NB], const T* const (&b)[NB],
                                             int D, int t, float (&out)[NB]) {
  constexpr int CH = RowChunk<MC>::N;
  float p[NB][4];
#pragma unroll
  for (int k = 0; k < NB; ++k) p[k][0] = p[k][1] = p[k][2] = p[k][3] = 0.f;
#pragma unroll
  for (int m0 = 0; m0 < MC; m0 += CH) {
    if (m0 * 64 >= D) break;
    float4 x[NB][CH], y[NB][CH];
    if ((m0 + CH) * 64 <= D) {
#pragma unroll
      for (int k = 0; k < NB; ++k)
#pragma unroll
        for (int c = 0; c < CH; ++c) {
          const int i = (m0 + c) * 64 + 4 * t;
          x[k][c] = load4(a[k] + i);
          y[k][c] = load4(b[k] + i);
        }
#pragma unroll
      for (int k = 0; k < NB; ++k)
#pragma unroll
        for (int c = 0; c < CH; ++c) fold4(x[k][c], y[k][c], p[k]);
    } else {
#pragma unroll
      for (int k = 0; k < NB; ++k)
#pragma unroll
        for (int c = 0; c < CH; ++c) {
          x[k][c] = load_chunk_tail(a[k], m0 + c, D, t);
          y[k][c] = load_chunk_tail(b[k], m0 + c, D, t);
        }
#pragma unroll
      for (int k = 0; k < NB; ++k)
#pragma unroll
        for (int c = 0; c < CH; ++c)
          if ((m0 + c) * 64 < D) fold4(x[k][c], y[k][c], p[k]);
    }
  }
#pragma unroll
  for (int k = 0; k < NB; ++k) out[k] = tree16(p[k][0], p[k][1], p[k][2], p[k][3]);
}

template <int MC, class T>
__device__ __forceinline__ float dist16(const Row16<MC>& a, const T* __restrict__ b, int D,
                                        int t) {
  float p0 = 0.f, p1 = 0.f, p2 = 0.f, p3 = 0.f;
#pragma unroll
  for (int m = 0; m < MC; ++m) {
    const int i = m * 64 + 4 * t;
    if (i < D) {
      const float4 y = load4(b + i);
      const float e0 = a.v[m].x - y.x, e1 = a.v[m].y - y.y, e2 = a.v[m].z - y.z,
                  e3 = a.v[m].w - y.w;
      p0 = __builtin_fmaf(e0, e0, p0);
      p1 = __builtin_fmaf(e1, e1, p1);
      p2 = __builtin_fmaf(e2, e2, p2);
      p3 = __builtin_fmaf(e3, e3, p3);
    }
  }
  return tree16(p0, p1, p2, p3);
}

// both rows from memory
template <int MC, class T>
__device__ __forceinline__ float dist16_mm(const T* __restrict__ a, const T* __restrict__ b,
                                           int D, int t) {
  float p0 = 0.f, p1 = 0.f, p2 = 0.f, p3 = 0.f;
#pragma unroll
  for (int m = 0; m < MC; ++m) {
    const int i = m * 64 + 4 * t;
    if (i < D) {
      const float4 x = load4(a + i);
      const float4 y = load4(b + i);
      const float e0 = x.x - y.x, e1 = x.y - y.y, e2 = x.z - y.z, e3 = x.w - y.w;
      p0 = __builtin_fmaf(e0, e0, p0);
      p1 = __builtin_fmaf(e1, e1, p1);
      p2 = __builtin_fmaf(e2, e2, p2);
      p3 = __builtin_fmaf(e3, e3, p3);
    }
  }
  return tree16(p0, p1, p2, p3);
}

// ---------------------------------------------------------------------------------------
// Filter rules (Tagore_src.cu:1488-1495).  res = d(p,c), dis = d(u,c), cur = d(u,p).
// Compiled with -ffp-contract=off and correctly rounded division / square root.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ float rule_eval(int rule, float res, float dis, float cur, float th) {
  if (rule == 0) return th * res - dis;
  return th - ((dis + cur) - res) / 2.0f / __builtin_sqrtf(dis * cur);
}

// ---------------------------------------------------------------------------------------
// Register sorts.  A wave holds 64*NPL keys; key e = r*64 + lane sits in slot r of lane
// `lane`.  A compare-exchange step with distance j < 64 takes the partner's key from lane
// lane ^ j: DPP for j = 1, 2 (quad permutations), 4 (row shifts by 4, selected per 4-lane
// bank) and 8 (row rotation by 8), ds_swizzle for j = 16 and ds_bpermute for j = 32; steps
// with j >= 64 swap slots inside the lane.  No LDS storage and no barriers.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lane_xor32(uint32_t v, int j) {
  const int x = int(v);
  switch (j) {
    case 1: return uint32_t(__builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, false));
    case 2: return uint32_t(__builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, false));
    case 4: {
      // banks 0 and 2 of each 16-lane row read lane + 4, banks 1 and 3 read lane - 4 (the
      // second move fills the lanes the first leaves unwritten)
      const int lo = __builtin_amdgcn_mov_dpp(x, 0x104, 0xF, 0x5, false);
      return uint32_t(__builtin_amdgcn_update_dpp(lo, x, 0x114, 0xF, 0xA, false));
    }
    case 8: return uint32_t(__builtin_amdgcn_mov_dpp(x, 0x128, 0xF, 0xF, false));
    case 16: return uint32_t(__builtin_amdgcn_ds_swizzle(x, 0x401F));  // xor 16 within 32
    default: return uint32_t(__builtin_amdgcn_ds_bpermute(int((lane_id() ^ j) << 2), x));
  }
}
__device__ __forceinline__ uint64_t lane_xor64(uint64_t v, int j) {
  return (uint64_t(lane_xor32(uint32_t(v >> 32), j)) << 32) | lane_xor32(uint32_t(v), j);
}
__device__ __forceinline__ uint32_t lane_xor(uint32_t v, int j) { return lane_xor32(v, j); }
__device__ __forceinline__ uint64_t lane_xor(uint64_t v, int j) { return lane_xor64(v, j); }

// The lanes whose index has bit b set (b < 6), as a wave mask.
__device__ __forceinline__ constexpr uint64_t lanes_with_bit(int b) {
  return b == 0   ? 0xAAAAAAAAAAAAAAAAull
         : b == 1 ? 0xCCCCCCCCCCCCCCCCull
         : b == 2 ? 0xF0F0F0F0F0F0F0F0ull
         : b == 3 ? 0xFF00FF00FF00FF00ull
         : b == 4 ? 0xFFFF0000FFFF0000ull
                  : 0xFFFFFFFF00000000ull;
}
// Whether a compare-exchange lane takes its partner's key p over its own x: p < x, negated
// in the lanes of `flip`.  The direction of every step depends only on lane bits, so flip is
// a constant wave mask: one 64-bit compare, one scalar XOR with the constant and the selects
// (a mask derived from the lane index at run time would be held in scalar registers across
// the whole sort, and spilled where they run short).
template <class T>
__device__ __forceinline__ bool take_partner(T p, T x, uint64_t flip) {
  return __builtin_amdgcn_inverse_ballot_w64(__builtin_amdgcn_ballot_w64(p < x) ^ flip);
}

// Bitonic sort of the 64*NPL keys v (layout above), ascending (descending with DESC).  A
// step costs one 64-bit compare and two selects per key besides the exchange.
template <int NPL, class T, bool DESC = false>
__device__ __forceinline__ void reg_bitonic(T (&v)[NPL]) {
  constexpr int N = 64 * NPL;
#pragma unroll
  for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (j >= 64) {
        const int rj = j >> 6;
#pragma unroll
        for (int r = 0; r < NPL; ++r) {
          if (r & rj) continue;
          const int r2 = r | rj;
          const bool asc = (((r << 6) & k) == 0) != DESC;  // k >= 128 here
          const T x = v[r], y = v[r2];
          const bool sw = (y < x) == asc;
          v[r] = sw ? y : x;
          v[r2] = sw ? x : y;
        }
      } else {
        // the lower position of an ascending block keeps the smaller key: the lane takes
        // the smaller key where (lower position) == (ascending block), else the larger
        const uint64_t upper = lanes_with_bit(__builtin_ctz(unsigned(j)));
#pragma unroll
        for (int r = 0; r < NPL; ++r) {
          uint64_t desc;  // lanes in a descending block
          if (k >= 64) desc = ((((r << 6) & k) == 0) != DESC) ? 0ull : ~0ull;
          else desc = lanes_with_bit(__builtin_ctz(unsigned(k))) ^ (DESC ? ~0ull : 0ull);
          const T x = v[r], p = lane_xor(x, j);
          v[r] = take_partner(p, x, upper ^ desc) ? p : x;
        }
      }
    }
  }
}

// One key per lane: sorts each block of KMAX lanes (KMAX a power of two up to 64) ascending.
// When only the first KMAX lanes hold keys below the rest, the whole wave ends up sorted.
template <int KMAX, class T>
__device__ __forceinline__ void reg_bitonic_upto(T& v) {
#pragma unroll
  for (int k = 2; k <= KMAX; k <<= 1) {
    const uint64_t desc = k == KMAX ? 0ull : lanes_with_bit(__builtin_ctz(unsigned(k)));
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      const T p = lane_xor(v, j);
      v = take_partner(p, v, lanes_with_bit(__builtin_ctz(unsigned(j))) ^ desc) ? p : v;
    }
  }
}

// Sorts the bitonic sequence of the 64*NPL keys v (layout above) ascending.
template <int NPL, class T>
__device__ __forceinline__ void reg_bitonic_merge(T (&v)[NPL]) {
#pragma unroll
  for (int j = 32 * NPL; j > 0; j >>= 1) {
    if (j >= 64) {
      const int rj = j >> 6;
#pragma unroll
      for (int r = 0; r < NPL; ++r) {
        if (r & rj) continue;
        const T x = v[r], y = v[r | rj];
        const bool sw = y < x;
        v[r] = sw ? y : x;
        v[r | rj] = sw ? x : y;
      }
    } else {
      const uint64_t upper = lanes_with_bit(__builtin_ctz(unsigned(j)));
#pragma unroll
      for (int r = 0; r < NPL; ++r) {
        const T x = v[r], p = lane_xor(x, j);
        v[r] = take_partner(p, x, upper) ? p : x;
      }
    }
  }
}
// reg_bitonic_merge with a 32-bit payload moved along with each key
template <int NPL>
__device__ __forceinline__ void reg_bitonic_merge_kv(uint64_t (&v)[NPL], uint32_t (&f)[NPL]) {
#pragma unroll
  for (int j = 32 * NPL; j > 0; j >>= 1) {
    if (j >= 64) {
      const int rj = j >> 6;
#pragma unroll
      for (int r = 0; r < NPL; ++r) {
        if (r & rj) continue;
        const bool sw = v[r | rj] < v[r];
        const uint64_t x = v[r], y = v[r | rj];
        const uint32_t fx = f[r], fy = f[r | rj];
        v[r] = sw ? y : x;
        v[r | rj] = sw ? x : y;
        f[r] = sw ? fy : fx;
        f[r | rj] = sw ? fx : fy;
      }
    } else {
      const uint64_t upper = lanes_with_bit(__builtin_ctz(unsigned(j)));
#pragma unroll
      for (int r = 0; r < NPL; ++r) {
        const uint64_t x = v[r], p = lane_xor64(x, j);
        const uint32_t pf = lane_xor32(f[r], j);
        // lower lanes take p < x, upper ones x < p: equal keys keep their own payload
        const bool take = __builtin_amdgcn_inverse_ballot_w64(
            (__builtin_amdgcn_ballot_w64(p < x) & ~upper) | (__builtin_amdgcn_ballot_w64(x < p) & upper));
        v[r] = take ? p : x;
        f[r] = take ? pf : f[r];
      }
    }
  }
}
__device__ __forceinline__ void reg_bitonic_merge64(uint64_t& v) {
  uint64_t w[1] = {v};
  reg_bitonic_merge<1>(w);
  v = w[0];
}

// ---------------------------------------------------------------------------------------
// Wave-level sorts in LDS.  Called by all 64 lanes of a one-wave workgroup; n is a power of
// two.  __syncthreads() orders the LDS traffic between network stages.
// ---------------------------------------------------------------------------------------
template <class T>
__device__ void wave_bitonic(T* a, int n) {
  const int lane = threadIdx.x;
  for (int k = 2; k <= n; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      // j is a power of two: 2 j (i / j) + i % j by shifts and a mask
      const int sh = __builtin_ctz(unsigned(j));
      for (int i = lane; i < (n >> 1); i += TG_WAVE) {
        const int lo = ((i >> sh) << (sh + 1)) | (i & (j - 1));
        const int hi = lo + j;
        const bool up = (lo & k) == 0;
        T x = a[lo], y = a[hi];
        if ((x > y) == up) {
          a[lo] = y;
          a[hi] = x;
        }
      }
      __syncthreads();
    }
  }
}

// In-place ascending sort of a[0..n) (n a power of two) by a one-wave workgroup: the keys
// go through registers (reg_bitonic, 64*NPL >= n slots, padding ~0 sorts last).  Larger n,
// or workgroups of several waves, use the LDS network above.  MAXN (<= 512) bounds the
// register variants compiled at a call site.
template <int NPL, class T>
__device__ __forceinline__ void reg_sort_lds(T* a, int n) {
  const int lane = lane_id();
  T v[NPL];
#pragma unroll
  for (int r = 0; r < NPL; ++r) v[r] = r * 64 + lane < n ? a[r * 64 + lane] : ~T(0);
  reg_bitonic<NPL>(v);
#pragma unroll
  for (int r = 0; r < NPL; ++r)
    if (r * 64 + lane < n) a[r * 64 + lane] = v[r];
}
template <int MAXN, class T>
__device__ void wave_sort(T* a, int n) {
  static_assert(MAXN <= 512, "register sorts hold at most 512 keys");
  if (blockDim.x != TG_WAVE || n > MAXN) {
    wave_bitonic(a, n);
    return;
  }
  if (n <= 64) reg_sort_lds<1>(a, n);
  else if (MAXN >= 128 && n <= 128) reg_sort_lds<(MAXN >= 128 ? 2 : 1)>(a, n);
  else if (MAXN >= 256 && n <= 256) reg_sort_lds<(MAXN >= 256 ? 4 : 1)>(a, n);
  else reg_sort_lds<(MAXN >= 512 ? 8 : 1)>(a, n);
  __syncthreads();
}

__host__ __device__ __forceinline__ int pow2_at_least(int n) {
  int p = 1;
  while (p < n) p <<= 1;
  return p;
}

// inclusive wave scan of an int
__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = lane_id();
#pragma unroll
  for (int s = 1; s < TG_WAVE; s <<= 1) {
    int o = __shfl_up(v, s, TG_WAVE);
    if (lane >= s) v += o;
  }
  return v;
}

// ---------------------------------------------------------------------------------------
// Push buffers ("update lists").  Producers append keys for a destination node: the first
// `cap` go to the node's slot row, the rest to a global overflow array of (dest, key)
// pairs.  After the producer kernel, overflow pairs are grouped per destination (count ->
// scan -> scatter, host-driven).  Reducers read slot row + overflow segment and sort, so
// the outcome is independent of arrival order.  Overflow beyond ov_cap sets *err.
// ---------------------------------------------------------------------------------------
struct PushBuf {
  uint32_t* cnt;       // [N] pushes per destination
  uint64_t* slots;     // [N * cap]
  uint32_t* ov_dest;   // [ov_cap]
  uint64_t* ov_key;    // [ov_cap]
  uint32_t* ov_used;   // [1]
  uint32_t* ov_off;    // [N + 1] exclusive offsets of grouped overflow
  uint64_t* ov_sorted; // [ov_cap] overflow keys grouped by destination
  uint32_t* err;       // [1]
  uint32_t cap;
  uint32_t ov_cap;
};

// Stores key at position s of v's push list (s from the list's counter): a slot, or the
// overflow area when the slots are full.
__device__ __forceinline__ void pb_put(const PushBuf& pb, uint32_t v, uint32_t s, uint64_t key) {
  if (s < pb.cap) {
    pb.slots[size_t(v) * pb.cap + s] = key;
  } else {
    // overflow entries of the active lanes take one allocation from the shared counter
    const uint64_t m = __ballot(1);
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (lane_id() == leader) base = atomicAdd(pb.ov_used, uint32_t(__popcll(m)));
    const uint32_t o = uint32_t(__shfl(int(base), leader, TG_WAVE)) + uint32_t(mask_rank(m));
    if (o < pb.ov_cap) {
      pb.ov_dest[o] = v;
      pb.ov_key[o] = key;
    } else {
      atomicOr(pb.err, 1u);
    }
  }
}

__device__ __forceinline__ void pb_push(const PushBuf& pb, uint32_t v, uint64_t key) {
  pb_put(pb, v, atomicAdd(&pb.cnt[v], 1u), key);
}

// Number of keys pushed to v, and accessor of the i-th (slot area first, then overflow).
__device__ __forceinline__ uint32_t pb_count(const PushBuf& pb, uint32_t v) { return pb.cnt[v]; }
__device__ __forceinline__ uint64_t pb_get(const PushBuf& pb, uint32_t v, uint32_t i) {
  if (i < pb.cap) return pb.slots[size_t(v) * pb.cap + i];
  const uint32_t o = pb.ov_off[v] + (i - pb.cap);
  // past ov_cap only when the overflow area itself overflowed (*err set, call fails)
  return o < pb.ov_cap ? pb.ov_sorted[o] : TG_KEY_INF;
}

// Keeps the R smallest keys pushed to v that pass `pred` (and, if uniq, merges identical
// keys), sorted ascending into buf[0..R); returns how many there are (<= R <= 128).
// buf must hold 256 u64 (LDS).  Optional seed keys (sorted, unique, <= R, all passing pred)
// start the running selection.
template <class Pred>
__device__ int pb_top(const PushBuf& pb, uint32_t v, int R, bool uniq, Pred pred, uint64_t* buf,
                      const uint64_t* seed = nullptr, int nseed = 0) {
  const int lane = threadIdx.x;
  const uint32_t n = pb_count(pb, v);
  if (uint32_t(nseed) + n <= 256) {
    // seed and pushes fit buf together (most nodes): one sort of the next power of two
    // above their count, then the first R (unique if requested) -- what the chunked path
    // selects
    const int tot = nseed + int(n);
    const int m = pow2_at_least(max(tot, 1));
    for (int i = lane; i < m; i += TG_WAVE) {
      uint64_t k = TG_KEY_INF;
      if (i < nseed) {
        k = seed[i];
      } else if (i < tot) {
        k = pb_get(pb, v, uint32_t(i - nseed));
        if (!pred(k)) k = TG_KEY_INF;
      }
      buf[i] = k;
    }
    __syncthreads();
    wave_sort<256>(buf, m);
    uint64_t kk[4];
    bool kp[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {  // every read before any write
      const int i = r * TG_WAVE + lane;
      kk[r] = i < m ? buf[i] : TG_KEY_INF;
      kp[r] = kk[r] != TG_KEY_INF && (!uniq || i == 0 || buf[i - 1] != kk[r]);
    }
    __syncthreads();
    int pos = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const uint64_t bm = ballot(kp[r]);
      const int p = pos + mask_rank(bm);
      if (kp[r] && p < R) buf[p] = kk[r];
      pos += int(__popcll(bm));
    }
    const int nb = pos < R ? pos : R;
    __syncthreads();
    for (int i = nb + lane; i < 256; i += TG_WAVE) buf[i] = TG_KEY_INF;
    __syncthreads();
    return nb;
  }
  int nbest = nseed;
  for (int i = lane; i < 128; i += TG_WAVE) buf[i] = i < nseed ? seed[i] : TG_KEY_INF;
  __syncthreads();
  for (uint32_t base = 0; base < n; base += 128) {
    for (int i = lane; i < 128; i += TG_WAVE) {
      uint64_t k = TG_KEY_INF;
      if (base + i < n) {
        k = pb_get(pb, v, base + i);
        if (!pred(k)) k = TG_KEY_INF;
      }
      buf[128 + i] = k;
    }
    __syncthreads();
    wave_sort<256>(buf, 256);
    // compaction: keep first R (unique if requested) into buf[0..R)
    uint64_t mine[4];
    bool keep[4];
    int cntk = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = lane * 4 + q;
      mine[q] = buf[i];
      keep[q] = mine[q] != TG_KEY_INF && (!uniq || i == 0 || buf[i - 1] != mine[q]);
      cntk += keep[q] ? 1 : 0;
    }
    const int incl = wave_incl_scan(cntk);
    const int total = __shfl(incl, TG_WAVE - 1, TG_WAVE);
    __syncthreads();
    int pos = incl - cntk;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (keep[q]) {
        if (pos < R) buf[pos] = mine[q];
        ++pos;
      }
    }
    nbest = total < R ? total : R;
    __syncthreads();
    for (int i = nbest + lane; i < 256; i += TG_WAVE) buf[i] = TG_KEY_INF;
    __syncthreads();
  }
  return nbest;
}

// lower_bound of key x in the sorted LDS array a[0..n)
template <class T>
__device__ __forceinline__ int lower_bound_lds(const T* a, int n, T x) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Is the key (d, id) an entry of kNN row v?  Rows are sorted by (distance, id) and hold
// canonical distances, so a listed id always appears under exactly this key.
__device__ __forceinline__ bool row_has_key(const uint32_t* G, const float* GD, int K, uint32_t v,
                                            uint64_t key) {
  const uint32_t* g = G + size_t(v) * K;
  const float* gd = GD + size_t(v) * K;
  int lo = 0, hi = K;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (mk_key(gd[mid], g[mid]) < key) lo = mid + 1; else hi = mid;
  }
  return lo < K && mk_key(gd[lo], g[lo]) == key;
}

template <class T>
__device__ __forceinline__ int upper_bound_lds(const T* a, int n, T x) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (!(x < a[mid])) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// dst[0..na+nb) = stable merge of the sorted arrays A[0..na) and B[0..nb) (A first on ties)
__device__ void merge_lds(const uint64_t* A, int na, const uint64_t* B, int nb, uint64_t* dst) {
  for (int i = threadIdx.x; i < na; i += TG_WAVE) dst[i + lower_bound_lds(B, nb, A[i])] = A[i];
  for (int j = threadIdx.x; j < nb; j += TG_WAVE) dst[j + upper_bound_lds(A, na, B[j])] = B[j];
}

// Loads of the u8 row at r in the quad layout (MC chunks; zero past D).
template <int MC>
__device__ __forceinline__ void load_quad_u8(uint4 (&y)[MC], const uint8_t* r, int D) {
  const int j = threadIdx.x & 3;
#pragma unroll
  for (int m = 0; m < MC; ++m) y[m] = m * 64 < D ? load_u8x16(r, m, D, j) : make_uint4(0u, 0u, 0u, 0u);
}
// |a - y|^2 for the register row a and the loaded row y (quad layout), in every quad lane.
template <int MC>
__device__ __forceinline__ float quad_dist_u8(const Row16<MC>& a, const uint4 (&y)[MC], int D) {
  uint32_t ab = 0, bb = 0;
#pragma unroll
  for (int m = 0; m < MC; ++m) {
    if (m * 64 >= D) break;
    ab = dot16_u8(as_u4(a.v[m]), y[m], ab);
    bb = dot16_u8(y[m], y[m], bb);
  }
  return float(a.n2 + quad_sum(int(bb) - 2 * int(ab)));
}

// d(ta, y) and d(tb, y) for the loaded row y (quad layout), in every quad lane.
template <int MC>
__device__ __forceinline__ void quad_dist2_u8(const Row16<MC>& ta, const Row16<MC>& tb,
                                              const uint4 (&y)[MC], int D, float& da, float& db) {
  uint32_t aa = 0, ab = 0, bb = 0;
#pragma unroll
  for (int m = 0; m < MC; ++m) {
    if (m * 64 >= D) break;
    aa = dot16_u8(as_u4(ta.v[m]), y[m], aa);
    ab = dot16_u8(as_u4(tb.v[m]), y[m], ab);
    bb = dot16_u8(y[m], y[m], bb);
  }
  da = float(ta.n2 + quad_sum(int(bb) - 2 * int(aa)));
  db = float(tb.n2 + quad_sum(int(bb) - 2 * int(ab)));
}

// Two u8 targets at once: each candidate row is read once and gives d(ta, c) and d(tb, c).
// One candidate per quad, 32 per step while more than 16 remain, then 16 (as eval_batch);
// emit(q, da, db) runs in lanes 0..31 (0..15) for the items q0 + lane.
template <int MC, class IdOf, class Emit>
__device__ __forceinline__ void eval_batch2_u8(const uint8_t* X, int D, const Row16<MC>& ta,
                                               const Row16<MC>& tb, int n, IdOf id_of,
                                               Emit emit) {
  const int lane = threadIdx.x, c = lane >> 2, src = (lane & 15) * 4;
  int q0 = 0;
  for (; q0 + 16 < n; q0 += 32) {
    const int qb = q0 + 16 + c;
    uint4 ya[MC], yb[MC];
    load_quad_u8<MC>(ya, X + size_t(id_of(q0 + c)) * D, D);  // q0 + c < q0 + 16 < n
    load_quad_u8<MC>(yb, qb < n ? X + size_t(id_of(qb)) * D : X, D);
    float da0, db0, da1, db1;
    quad_dist2_u8<MC>(ta, tb, ya, D, da0, db0);
    quad_dist2_u8<MC>(ta, tb, yb, D, da1, db1);
    const float xa0 = __shfl(da0, src, TG_WAVE), xb0 = __shfl(db0, src, TG_WAVE);
    const float xa1 = __shfl(da1, src, TG_WAVE), xb1 = __shfl(db1, src, TG_WAVE);
    if (lane < 32 && q0 + lane < n) emit(q0 + lane, lane < 16 ? xa0 : xa1, lane < 16 ? xb0 : xb1);
  }
  if (q0 < n) {
    const int q = q0 + c;
    uint4 y[MC];
    load_quad_u8<MC>(y, q < n ? X + size_t(id_of(q)) * D : X, D);  // unused quads read row 0
    float da, db;
    quad_dist2_u8<MC>(ta, tb, y, D, da, db);
    const float va = __shfl(da, src, TG_WAVE), vb = __shfl(db, src, TG_WAVE);
    if (lane < 16 && q0 + lane < n) emit(q0 + lane, va, vb);
  }
}

// Evaluates the targets id_of(0..n) against the register row `tgt`, emit(q, d) once per item.
//  * u8 rows (quad layout): one candidate per quad; with WIDE, 64 per step (items
//    q0 + 16h + lane/4, h < 4: four rows in flight per quad, 16 more VGPRs) while more than
//    48 remain; then 32 per step while more than 16 remain, then 16; the results are moved
//    to lanes 0..63 (0..31, 0..15), lane j holding item q0 + j, and emit runs there;
//  * FP32 / BF16 rows: four candidates per 16-lane group at a time (item q0 + g + 4b in group
//    g); the 16 results are moved to lanes 0..15 and emit runs there, once per item.
// Every emit writes only what its own item owns, so the lanes may run it together.
template <int MC, class T, bool WIDE = false, class IdOf, class Emit>
__device__ __forceinline__ void eval_batch(const T* X, int D, const Row16<MC>& tgt, int n,
                                           IdOf id_of, Emit emit) {
  const int lane = threadIdx.x;
  if constexpr (std::is_same<T, uint8_t>::value) {
    const int c = lane >> 2, src = (lane & 15) * 4;
    int q0 = 0;
    for (; WIDE && q0 + 48 < n; q0 += 64) {
      // 64 per step: four rows in flight per quad
      const int qb = q0 + 16 + c, qc = q0 + 32 + c, qd = q0 + 48 + c;
      uint4 ya[MC], yb[MC], yc[MC], yd[MC];
      load_quad_u8<MC>(ya, X + size_t(id_of(q0 + c)) * D, D);  // qc < q0 + 48 < n
      load_quad_u8<MC>(yb, X + size_t(id_of(qb)) * D, D);
      load_quad_u8<MC>(yc, X + size_t(id_of(qc)) * D, D);
      load_quad_u8<MC>(yd, qd < n ? X + size_t(id_of(qd)) * D : X, D);
      const float da = quad_dist_u8<MC>(tgt, ya, D), db = quad_dist_u8<MC>(tgt, yb, D);
      const float dc = quad_dist_u8<MC>(tgt, yc, D), dd = quad_dist_u8<MC>(tgt, yd, D);
      const float xa = __shfl(da, src, TG_WAVE), xb = __shfl(db, src, TG_WAVE);
      const float xc = __shfl(dc, src, TG_WAVE), xd = __shfl(dd, src, TG_WAVE);
      const float v = lane < 16 ? xa : lane < 32 ? xb : lane < 48 ? xc : xd;
      if (q0 + lane < n) emit(q0 + lane, v);
    }
    for (; q0 + 16 < n; q0 += 32) {
      const int qb = q0 + 16 + c;
      uint4 ya[MC], yb[MC];
      load_quad_u8<MC>(ya, X + size_t(id_of(q0 + c)) * D, D);  // q0 + c < q0 + 16 < n
      load_quad_u8<MC>(yb, qb < n ? X + size_t(id_of(qb)) * D : X, D);
      const float da = quad_dist_u8<MC>(tgt, ya, D), db = quad_dist_u8<MC>(tgt, yb, D);
      const float xa = __shfl(da, src, TG_WAVE), xb = __shfl(db, src, TG_WAVE);
      const float v = lane < 16 ? xa : xb;
      if (lane < 32 && q0 + lane < n) emit(q0 + lane, v);
    }
    if (q0 < n) {
      const int q = q0 + c;
      uint4 y[MC];
      load_quad_u8<MC>(y, q < n ? X + size_t(id_of(q)) * D : X, D);
      const float d = quad_dist_u8<MC>(tgt, y, D);
      const float v = __shfl(d, src, TG_WAVE);
      if (lane < 16 && q0 + lane < n) emit(q0 + lane, v);
    }
  } else {
    const int g = lane >> 4, t = lane & 15;
    const int src = 16 * (lane & 3), mine = (lane >> 2) & 3;
    for (int q0 = 0; q0 < n; q0 += 16) {
      const T* rows[4];
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int q = q0 + g + 4 * b;
        rows[b] = q < n ? X + size_t(id_of(q)) * D : X;  // unused slots read row 0
      }
      float dd[4];
      dist16_batch<MC, 4>(tgt, rows, D, t, dd);
      float v = 0.f;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const float x = __shfl(dd[b], src, TG_WAVE);
        if (b == mine) v = x;
      }
      if (lane < 16 && q0 + lane < n) emit(q0 + lane, v);
    }
  }
}

}  // namespace tg
