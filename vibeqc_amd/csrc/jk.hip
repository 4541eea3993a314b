// J and K contraction kernels for gfx950 (see jk.hpp for the data layout).
//
// MFMA usage:
//   FP64   v_mfma_f64_16x16x4_f64   (A/B one element per lane, K = 4)
//   FP32   v_mfma_f32_16x16x4_f32   (same operand layout, FP32 accumulation)
//   BF16X3 v_mfma_f32_16x16x32_bf16 (8 bf16 per lane, K = 32); each FP32
//          operand x is split as hi = bf16(x), lo = bf16(x - hi) and the
//          product is accumulated as hi*hi + hi*lo + lo*hi in FP32.
// Accumulator layout (measured on gfx950): column lane % 16 in every case;
// row 4*(lane/16) + r in register r for the FP32 and BF16 instructions, but
// row lane/16 + 4*r for v_mfma_f64_16x16x4_f64 (AccRow below).
#include "jk.hpp"

#include <algorithm>
#include <cmath>

namespace vqc {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
typedef float fl2 __attribute__((ext_vector_type(2)));

template <class T>
struct AccOf;
template <>
struct AccOf<double> {
  using type = d4;
};
template <>
struct AccOf<float> {
  using type = f4;
};

// Row of accumulator register r held by `lane` in a 16x16 MFMA tile.
template <class T>
__device__ __forceinline__ int acc_row(int lane, int r) {
  return 4 * (lane >> 4) + r;
}
template <>
__device__ __forceinline__ int acc_row<double>(int lane, int r) {
  return (lane >> 4) + 4 * r;
}

__device__ __forceinline__ d4 mfma(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4 mfma(float a, float b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4 mfma(bf8 a, bf8 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ std::size_t packed_index(int mu, int nu) {
  const std::size_t hi = mu > nu ? mu : nu, lo = mu > nu ? nu : mu;
  return hi * (hi + 1) / 2 + lo;
}

__device__ __forceinline__ void unpack_pair(std::size_t p, int& mu, int& nu) {
  int m = static_cast<int>((sqrt(8.0 * static_cast<double>(p) + 1.0) - 1.0) * 0.5);
  while (static_cast<std::size_t>(m) * (m + 1) / 2 > p) --m;
  while (static_cast<std::size_t>(m + 1) * (m + 2) / 2 <= p) ++m;
  mu = m;
  nu = static_cast<int>(p - static_cast<std::size_t>(m) * (m + 1) / 2);
}

__device__ __forceinline__ void split_bf16(float x, __bf16& hi, __bf16& lo) {
  hi = static_cast<__bf16>(x);
  lo = static_cast<__bf16>(x - static_cast<float>(hi));
}

// LDS tiles of 32 bf16 per row (64 bytes, four 16-byte chunks, no padding)
// that MFMA fragments read as 16 consecutive rows at one chunk: chunk c of
// row r is stored at position c ^ ((r >> 1) & 3), so the 16 reads of a
// fragment cover all 64 banks once (a padded 80-byte row leaves 2-way
// conflicts; measured 3.4e9 conflict cycles per (H2O)64 U launch, zero with
// this layout).
__device__ __forceinline__ int lds_chunk(int row, int chunk) { return chunk ^ ((row >> 1) & 3); }

// Split of eight floats into bf16 hi/lo fragments, two at a time (one packed
// conversion per pair for hi and one for lo).
__device__ __forceinline__ void split8_bf16(const float (&x)[8], bf8& hi, bf8& lo) {
  bf2 h[4], l[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const fl2 v = {x[2 * k], x[2 * k + 1]};
    h[k] = __builtin_convertvector(v, bf2);
    l[k] = __builtin_convertvector(v - __builtin_convertvector(h[k], fl2), bf2);
  }
  hi = __builtin_shufflevector(__builtin_shufflevector(h[0], h[1], 0, 1, 2, 3), __builtin_shufflevector(h[2], h[3], 0, 1, 2, 3),
                               0, 1, 2, 3, 4, 5, 6, 7);
  lo = __builtin_shufflevector(__builtin_shufflevector(l[0], l[1], 0, 1, 2, 3), __builtin_shufflevector(l[2], l[3], 0, 1, 2, 3),
                               0, 1, 2, 3, 4, 5, 6, 7);
}

// ---------------------------------------------------------------------------
// J
// ---------------------------------------------------------------------------

__global__ void pack_density_kernel(const double* __restrict__ D, int n, double* __restrict__ dp) {
  const std::size_t p = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= static_cast<std::size_t>(n) * (n + 1) / 2) return;
  int mu, nu;
  unpack_pair(p, mu, nu);
  const std::size_t a = static_cast<std::size_t>(mu) * n + nu, b = static_cast<std::size_t>(nu) * n + mu;
  dp[p] = mu == nu ? D[a] : D[a] + D[b];
}

constexpr int kRhoThreads = 256;

template <class T, int V>
__global__ __launch_bounds__(kRhoThreads) void coulomb_rho_kernel(const T* __restrict__ B,
                                                                  const double* __restrict__ dp,
                                                                  std::size_t npair, int ldq, std::size_t chunk,
                                                                  double* __restrict__ partial) {
  const int q = (blockIdx.x * kRhoThreads + threadIdx.x) * V;
  if (q >= ldq) return;
  const std::size_t p0 = static_cast<std::size_t>(blockIdx.y) * chunk;
  const std::size_t p1 = p0 + chunk < npair ? p0 + chunk : npair;
  double acc[V];
#pragma unroll
  for (int v = 0; v < V; ++v) acc[v] = 0.0;
  for (std::size_t p = p0; p < p1; ++p) {
    const double w = dp[p];
    const T* row = B + p * ldq + q;
#pragma unroll
    for (int v = 0; v < V; ++v) acc[v] += w * static_cast<double>(row[v]);
  }
  double* out = partial + static_cast<std::size_t>(blockIdx.y) * ldq + q;
#pragma unroll
  for (int v = 0; v < V; ++v) out[v] = acc[v];
}

// out[c] = sum_r partial[r][c]: 64 columns per block, 16 row groups per
// column reduced through LDS in a fixed order (deterministic).
constexpr int kSumCols = 64, kSumGroups = 16;
__global__ __launch_bounds__(kSumCols * kSumGroups) void sum_rows_kernel(const double* __restrict__ partial, int rows,
                                                                        int cols, double* __restrict__ out) {
  __shared__ double red[kSumGroups][kSumCols];
  const int lane = threadIdx.x % kSumCols, g = threadIdx.x / kSumCols;
  const int c = blockIdx.x * kSumCols + lane;
  double s = 0.0;
  if (c < cols) {
#pragma unroll 4
    for (int r = g; r < rows; r += kSumGroups) s += partial[static_cast<std::size_t>(r) * cols + c];
  }
  red[g][lane] = s;
  __syncthreads();
  if (g == 0 && c < cols) {
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < kSumGroups; ++k) t += red[k][lane];
    out[c] = t;
  }
}

template <class T>
struct Vec16;
template <>
struct Vec16<double> {
  using type = double __attribute__((ext_vector_type(2)));
};
template <>
struct Vec16<float> {
  using type = float __attribute__((ext_vector_type(4)));
};

constexpr int kJThreads = 1024, kJInFlight = 12;

// J_p = sum_Q B[p][Q] rho_Q, one wavefront per packed row.  Each lane keeps
// kJInFlight 16-byte loads of its row in flight before consuming them (12
// cover a 2432-wide FP32 row in one batch), and a 1024-thread block shares
// one LDS copy of rho, so with up to 128 VGPRs per lane one block (16 waves)
// per CU keeps enough bytes outstanding to stream B at HBM rate.
template <class T>
__global__ __launch_bounds__(kJThreads) __attribute__((amdgpu_waves_per_eu(4))) void coulomb_matrix_kernel(const T* __restrict__ B,
                                                                  const double* __restrict__ rho,
                                                                  std::size_t npair, int ldq, int n,
                                                                  double* __restrict__ J) {
  using V = typename Vec16<T>::type;
  constexpr int W = 16 / sizeof(T), STEP = 64 * W;
  extern __shared__ double srho[];
  for (int q = threadIdx.x; q < ldq; q += blockDim.x) srho[q] = rho[q];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const std::size_t wave = static_cast<std::size_t>(blockIdx.x) * (blockDim.x / 64) + (threadIdx.x >> 6);
  const std::size_t nwave = static_cast<std::size_t>(gridDim.x) * (blockDim.x / 64);
  for (std::size_t p = wave; p < npair; p += nwave) {
    const T* row = B + p * ldq;
    double s = 0.0;
    for (int q0 = lane * W; q0 < ldq; q0 += STEP * kJInFlight) {
      V v[kJInFlight];
#pragma unroll
      for (int u = 0; u < kJInFlight; ++u) {
        const int q = q0 + u * STEP;
        v[u] = q < ldq ? *reinterpret_cast<const V*>(row + q) : V{};
      }
#pragma unroll
      for (int u = 0; u < kJInFlight; ++u) {
        const int q = q0 + u * STEP;
        if (q < ldq) {
#pragma unroll
          for (int e = 0; e < W; ++e) s += static_cast<double>(v[u][e]) * srho[q + e];
        }
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) {
      int mu, nu;
      unpack_pair(p, mu, nu);
      J[static_cast<std::size_t>(mu) * n + nu] = s;
      J[static_cast<std::size_t>(nu) * n + mu] = s;
    }
  }
}

// ---------------------------------------------------------------------------
// K, step 1: U[mu][i][Q] = sum_nu B[(mu,nu)][Q] C[nu][i]
// Block: 2 x WQ waves (WQ along Q x 2 along i); tile 32*WQ (Q) x 32*NIW (i),
// i.e. all occupied columns up to 320 in one block, so each packed row of B
// is streamed once per mu.  Every wave owns 2 x NIW MFMA 16x16 tiles.  FP32 and
// BF16X3 use WQ = 8 (1024 threads, 256 rows of Q per C tile); FP64 uses
// WQ = 4 so its 2 x NIW x 4 double accumulators fit the register budget.
// ---------------------------------------------------------------------------

constexpr int kUTileQ = 256;

template <class T, int NIW, int WQ>
__global__ __launch_bounds__(128 * WQ) void occupied_u_kernel(const T* __restrict__ B, const T* __restrict__ C,
                                                             T* __restrict__ U, std::size_t us, int n, int ldq,
                                                             int ldc, int r, int nqb, const double* __restrict__ D,
                                                             double* __restrict__ rho_part) {
  using Acc = typename AccOf<T>::type;
  constexpr int BQ = 32 * WQ, BK = 16, BI = 32 * NIW, NT = 128 * WQ;
  constexpr int CP = (BK * BI + NT - 1) / NT;  // C elements staged per thread
  __shared__ T Bs[2][BK][BQ];
  __shared__ T Cs[2][BK][BI];
  // Q-tile-major order: the blocks in flight share one Q tile, so the
  // second gather of each packed row (as (mu, nu) and as (nu, mu)) finds it
  // in the last-level cache instead of HBM.
  const int mu = blockIdx.x % n;
  const int q0 = (blockIdx.x / n) * BQ;
  const int i0 = blockIdx.y * BI;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wq = wave % WQ, wi = wave / WQ;
  Acc acc[2][NIW];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < NIW; ++y) acc[x][y] = Acc{};
  const int bk = tid / (BQ / 4), bq = (tid % (BQ / 4)) * 4;
  // The first i-block also forms rho_Q partials, sum_nu D[mu][nu] B[(mu,nu)][Q],
  // from the B values it streams anyway (FP64 accumulation).
  const bool with_rho = rho_part != nullptr && blockIdx.y == 0;
  double racc[4] = {0.0, 0.0, 0.0, 0.0};
  extern __shared__ double drow[];  // D[mu][:], staged once when rho is formed
  if (with_rho)
    for (int k = tid; k < n; k += NT) drow[k] = D[static_cast<std::size_t>(mu) * n + k];
  __syncthreads();
  T rb[4], rc[CP];
  double rd = 0.0;
  // Stage the next k-step in registers so its global latency overlaps the
  // MFMAs of the current one; LDS is double-buffered (one barrier per step).
  // Staged values are only consumed in sstore, after the MFMAs.
  auto gload = [&](int nu0) {
    const int nu = nu0 + bk, q = q0 + bq;
    if (nu < n && q < ldq) {
      const T* src = B + packed_index(mu, nu) * ldq + q;
#pragma unroll
      for (int e = 0; e < 4; ++e) rb[e] = src[e];
      rd = nu;
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) rb[e] = T(0);
      rd = -1.0;
    }
#pragma unroll
    for (int j = 0; j < CP; ++j) {
      const int e = tid + j * NT;
      const int k = e / BI, c = e % BI, nu2 = nu0 + k;
      rc[j] = (e < BK * BI && nu2 < n) ? C[static_cast<std::size_t>(nu2) * ldc + i0 + c] : T(0);
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int e = 0; e < 4; ++e) Bs[buf][bk][bq + e] = rb[e];
    if (with_rho) {
      const double d = rd >= 0.0 ? drow[static_cast<int>(rd)] : 0.0;
#pragma unroll
      for (int e = 0; e < 4; ++e) racc[e] += d * static_cast<double>(rb[e]);
    }
#pragma unroll
    for (int j = 0; j < CP; ++j) {
      const int e = tid + j * NT;
      if (e < BK * BI) Cs[buf][e / BI][e % BI] = rc[j];
    }
  };
  gload(0);
  sstore(0);
  __syncthreads();
  int cur = 0;
  for (int nu0 = 0; nu0 < n; nu0 += BK) {
    const bool more = nu0 + BK < n;
    if (more) gload(nu0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int k = kk + (lane >> 4);
      const T a0 = Bs[cur][k][wq * 32 + (lane & 15)];
      const T a1 = Bs[cur][k][wq * 32 + 16 + (lane & 15)];
#pragma unroll
      for (int y = 0; y < NIW; ++y) {
        const T bv = Cs[cur][k][wi * 16 * NIW + y * 16 + (lane & 15)];
        acc[0][y] = mfma(a0, bv, acc[0][y]);
        acc[1][y] = mfma(a1, bv, acc[1][y]);
      }
    }
    if (more) sstore(cur ^ 1);
    __syncthreads();
    cur ^= 1;
  }
  if (with_rho) {
    // Reduce the BK row groups of each Q column through LDS (the B tiles are free).
    double* red = reinterpret_cast<double*>(&Bs[0][0][0]);
#pragma unroll
    for (int e = 0; e < 4; ++e) red[bk * BQ + bq + e] = racc[e];
    __syncthreads();
    if (tid < BQ && q0 + tid < ldq) {
      double t = 0.0;
#pragma unroll
      for (int k = 0; k < BK; ++k) t += red[k * BQ + tid];
      rho_part[static_cast<std::size_t>(mu) * ldq + q0 + tid] = t;
    }
  }
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < NIW; ++y) {
      const int i = i0 + wi * 16 * NIW + y * 16 + (lane & 15);
      const int qb = q0 + wq * 32 + x * 16;
      if (i >= r || qb >= ldq) continue;
      T* dst = U + static_cast<std::size_t>(mu) * us + static_cast<std::size_t>(i) * ldq + qb;
#pragma unroll
      for (int v = 0; v < 4; ++v) dst[acc_row<T>(lane, v)] = acc[x][y][v];
    }
}

// BF16X3 variant: B (packed FP32) is split into hi/lo bfloat16 and transposed
// into K-contiguous LDS rows: each thread gathers 8 consecutive nu of one Q
// (a wave reads 64 consecutive Q per row, coalesced) and writes one 16-byte
// chunk per plane.  C^T arrives pre-split (ldc rows of ldct K-contiguous
// entries, zero beyond n).  U is written as hi/lo planes.  The next k-step is
// staged in registers while the current one runs on the MFMAs.
template <int NIW, int WQ, int NBUF>
__global__ __launch_bounds__(1024) void occupied_u_bf16x3_kernel(const float* __restrict__ B,
                                                                const __bf16* __restrict__ Cth,
                                                                const __bf16* __restrict__ Ctl,
                                                                __bf16* __restrict__ Uh, __bf16* __restrict__ Ul,
                                                                std::size_t us, int n, int ldq, int ldct, int r,
                                                                int nqb,
                                                                const double* __restrict__ D,
                                                                double* __restrict__ rho_part) {
  // 16 waves: WQ along Q x (16 / WQ) along i.  WQ = 4 gives 64 x 8*NIW wave
  // tiles (even NIW), which read fewer LDS fragments per MFMA than WQ = 8
  // (32 x 16*NIW).
  constexpr int BQ = kUTileQ, BK = 32, BI = 32 * NIW, NT = 1024;
  constexpr int WI = 16 / WQ, XF = BQ / (16 * WQ), YF = BI / (16 * WI);
  static_assert(YF * 16 * WI == BI, "i tile must split evenly over the waves");
  constexpr int YP = (BI * 4 + NT - 1) / NT;  // 16-byte C chunks per thread and plane
  // NBUF = 2: the waves that finish the MFMAs of step k store step k + 1
  // into the other LDS buffer while the rest are still multiplying, with one
  // barrier per step.  NBUF = 1 (wide i tiles, whose two buffers would not
  // fit beside the rho row): store after a barrier, then a second barrier.
  // Rows of 64 bytes with swizzled chunks (lds_chunk).
  __shared__ __bf16 Xh[NBUF][BQ][BK], Xl[NBUF][BQ][BK];
  __shared__ __bf16 Yh[NBUF][BI][BK], Yl[NBUF][BI][BK];
  // Q-tile-major order: the blocks in flight share one Q tile, so the
  // second gather of each packed row (as (mu, nu) and as (nu, mu)) finds it
  // in the last-level cache instead of HBM.
  const int mu = blockIdx.x % n;
  const int q0 = (blockIdx.x / n) * BQ;
  const int i0 = blockIdx.y * BI;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wq = wave % WQ, wi = wave / WQ;
  f4 acc[XF][YF];
#pragma unroll
  for (int x = 0; x < XF; ++x)
#pragma unroll
    for (int y = 0; y < YF; ++y) acc[x][y] = f4{};
  // One Q column and 8 consecutive nu per thread.  BQ is a multiple of the
  // wave size, so the nu group is uniform across a wave: every gathered row
  // address is computed once per wave in scalar registers and each lane adds
  // only its Q offset.  Lanes past ldq (the last Q tile) read column 0 and
  // keep a zero.
  static_assert(BQ % 64 == 0, "a wave must stay inside one nu group");
  const int xq = tid % BQ;
  const int xg = __builtin_amdgcn_readfirstlane(tid / BQ);
  const bool q_ok = q0 + xq < ldq;
  const int q_off = q_ok ? q0 + xq : 0;
  const bool with_rho = rho_part != nullptr && blockIdx.y == 0;
  double racc = 0.0;
  extern __shared__ double drow[];  // D[mu][:], staged once when rho is formed
  if (with_rho)
    for (int k = tid; k < n; k += NT) drow[k] = D[static_cast<std::size_t>(mu) * n + k];
  __syncthreads();
  // Only the B gather (HBM) is staged in registers, one step ahead; the small
  // C^T tile is L2-resident and is copied straight into LDS.  The first
  // i-block also accumulates rho_Q partials from the staged B values (FP32
  // B, FP64 sums) when they are consumed in sstore, so the loads stay in
  // flight.
  // The eight rows of a step are gathered without branches, so all eight
  // loads are in flight together: the row pointer starts at nu = min(nb,
  // n - 1) and advances in scalar registers by one row below the diagonal
  // (rows (mu, nu) are consecutive) and by nu rows above it (row (nu, mu)
  // follows (nu - 1, mu) after nu rows); rows past n repeat the last valid
  // one and are zeroed when stored.
  float rx[8];
  auto gload = [&](int nu0) {
    const int nb = nu0 + xg * 8;
    const float* row = B + packed_index(mu, nb < n ? nb : n - 1) * ldq;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      rx[e] = row[q_off];
      if (e < 7) {
        // Branch-free step (a branch here would make the compiler wait for
        // each load at the join).
        const int next = nb + e + 1;
        const int step = next < n ? (next <= mu ? 1 : next) : 0;
        row += static_cast<std::size_t>(step) * ldq;
      }
    }
  };
  auto sstore = [&](int buf, int nu0) {
#pragma unroll
    for (int e = 0; e < 8; ++e)
      if (nu0 + xg * 8 + e >= n || !q_ok) rx[e] = 0.0f;
    if (with_rho) {
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int nu = nu0 + xg * 8 + e;
        if (nu < n) racc += drow[nu] * static_cast<double>(rx[e]);
      }
    }
    bf8 h, l;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      __bf16 hh, ll;
      split_bf16(rx[e], hh, ll);
      h[e] = hh;
      l[e] = ll;
    }
    *reinterpret_cast<bf8*>(&Xh[buf][xq][lds_chunk(xq, xg) * 8]) = h;
    *reinterpret_cast<bf8*>(&Xl[buf][xq][lds_chunk(xq, xg) * 8]) = l;
#pragma unroll
    for (int j = 0; j < YP; ++j) {
      const int e = tid + j * NT;
      if (e < BI * 4) {
        const std::size_t off = static_cast<std::size_t>(i0 + (e >> 2)) * ldct + nu0 + (e & 3) * 8;
        const int yc = lds_chunk(e >> 2, e & 3) * 8;
        *reinterpret_cast<bf8*>(&Yh[buf][e >> 2][yc]) = *reinterpret_cast<const bf8*>(Cth + off);
        *reinterpret_cast<bf8*>(&Yl[buf][e >> 2][yc]) = *reinterpret_cast<const bf8*>(Ctl + off);
      }
    }
  };
  gload(0);
  sstore(0, 0);
  if (BK < n) gload(BK);
  __syncthreads();
  int cur = 0;
  for (int nu0 = 0; nu0 < n; nu0 += BK) {
    // Fragment rows start at multiples of 16, so the swizzled chunk depends
    // on the lane only.
    const int kq = lds_chunk(lane & 15, lane >> 4) * 8;
    bf8 ah[XF], al[XF];
#pragma unroll
    for (int x = 0; x < XF; ++x) {
      ah[x] = *reinterpret_cast<const bf8*>(&Xh[cur][wq * 16 * XF + x * 16 + (lane & 15)][kq]);
      al[x] = *reinterpret_cast<const bf8*>(&Xl[cur][wq * 16 * XF + x * 16 + (lane & 15)][kq]);
    }
#pragma unroll
    for (int y = 0; y < YF; ++y) {
      const int row = wi * 16 * YF + y * 16 + (lane & 15);
      const bf8 bh = *reinterpret_cast<const bf8*>(&Yh[cur][row][kq]);
      const bf8 bl = *reinterpret_cast<const bf8*>(&Yl[cur][row][kq]);
      // XF independent accumulators between dependent MFMAs keep the pipe full.
#pragma unroll
      for (int x = 0; x < XF; ++x) acc[x][y] = mfma(ah[x], bh, acc[x][y]);
#pragma unroll
      for (int x = 0; x < XF; ++x) acc[x][y] = mfma(ah[x], bl, acc[x][y]);
#pragma unroll
      for (int x = 0; x < XF; ++x) acc[x][y] = mfma(al[x], bh, acc[x][y]);
    }
    if constexpr (NBUF == 2) {
      if (nu0 + BK < n) {
        // The other buffer was last read before the previous barrier.
        sstore(cur ^ 1, nu0 + BK);
        if (nu0 + 2 * BK < n) gload(nu0 + 2 * BK);
      }
      __syncthreads();
      cur ^= 1;
    } else {
      __syncthreads();
      if (nu0 + BK < n) {
        sstore(0, nu0 + BK);
        if (nu0 + 2 * BK < n) gload(nu0 + 2 * BK);
        __syncthreads();
      }
    }
  }
  if (with_rho) {
    static_assert(sizeof(Xh) >= NT * sizeof(double), "the rho reduction reuses Xh");
    double* red = reinterpret_cast<double*>(&Xh[0][0][0]);
    red[xg * BQ + xq] = racc;
    __syncthreads();
    if (tid < BQ && q0 + tid < ldq) {
      double t = 0.0;
#pragma unroll
      for (int g = 0; g < NT / BQ; ++g) t += red[g * BQ + tid];
      rho_part[static_cast<std::size_t>(mu) * ldq + q0 + tid] = t;
    }
  }
#pragma unroll
  for (int x = 0; x < XF; ++x)
#pragma unroll
    for (int y = 0; y < YF; ++y) {
      const int i = i0 + wi * 16 * YF + y * 16 + (lane & 15);
      const int qb = q0 + wq * 16 * XF + x * 16;
      if (i >= r || qb >= ldq) continue;
      const std::size_t off = static_cast<std::size_t>(mu) * us + static_cast<std::size_t>(i) * ldq + qb + 4 * (lane >> 4);
#pragma unroll
      for (int v = 0; v < 4; ++v) split_bf16(acc[x][y][v], Uh[off + v], Ul[off + v]);
    }
}

// BF16X3 U with the B operand taken straight from registers (i tiles up to
// 96 columns).  Lane l of a wave gathers the adjacent Q rows q + 2 (l & 15)
// and q + 2 (l & 15) + 1 (one 8-byte load) at the eight consecutive nu of k
// group l >> 4: the even rows form one A fragment of
// v_mfma_f32_16x16x32_bf16 and the odd rows another, each in exactly the
// operand layout, split into hi/lo in registers; B never passes through LDS.
// Only the C^T tile (i x nu) is staged in LDS, double-buffered, one barrier
// per k-step.  8 waves x 32 Q rows per block (two blocks per CU).  The C^T
// global loads of the next step are issued before its B gathers, so the wait
// that precedes their LDS store never waits for the gathers.
template <int NIW, int TERMS>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(NIW <= 3 ? 4 : 2))) void occupied_u_bf16x3_direct_kernel(
    const float* __restrict__ B, const __bf16* __restrict__ Cth, const __bf16* __restrict__ Ctl,
    __bf16* __restrict__ Uh, __bf16* __restrict__ Ul, std::size_t us, int n, int ldq, int ldct, int r,
    const double* __restrict__ D, double* __restrict__ rho_part) {
  constexpr int NW = 8, NT = 64 * NW, BQ = 32 * NW, BK = 32, BI = 32 * NIW;
  constexpr int XF = 2, YF = BI / 16;
  constexpr int YP = (BI * 4 + NT - 1) / NT;  // 16-byte C^T chunks per thread and plane
  // C^T rows of 64 bytes, chunks swizzled (lds_chunk): the sixteen rows one
  // fragment read touches hit distinct banks.
  __shared__ __bf16 Yh[2][BI][BK], Yl[2][BI][BK];
  // Dynamic LDS: element offsets of the packed rows (mu, nu) for nu < n + BK
  // (rows past n repeat row n - 1, so gathers need no bounds test), then, for
  // the rho pass, D[mu][nu] as float.
  extern __shared__ double dyn[];
  std::size_t* rowoff = reinterpret_cast<std::size_t*>(dyn);
  float* drow = reinterpret_cast<float*>(rowoff + n + BK);
  const int mu = blockIdx.x % n;  // Q-tile-major block order
  const int q0 = (blockIdx.x / n) * BQ;
  const int i0 = blockIdx.y * BI;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int kg = lane >> 4;  // k group: nu offset 8 kg within a step
  const int qw = q0 + wave * 32;
  const bool q_ok = qw < ldq;  // wave-uniform: ldq is a multiple of 32
  const int q_off = q_ok ? qw + 2 * (lane & 15) : 0;  // rows q_off (x = 0), q_off + 1 (x = 1)
  const bool with_rho = rho_part != nullptr && blockIdx.y == 0;
  for (int k = tid; k < n + BK; k += NT) rowoff[k] = packed_index(mu, k < n ? k : n - 1) * ldq;
  if (with_rho)
    for (int k = tid; k < n; k += NT) drow[k] = static_cast<float>(D[static_cast<std::size_t>(mu) * n + k]);
  f4 acc[XF][YF];
#pragma unroll
  for (int x = 0; x < XF; ++x)
#pragma unroll
    for (int y = 0; y < YF; ++y) acc[x][y] = f4{};
  double racc[XF] = {0.0, 0.0};
  fl2 rx[8] = {};
  bf8 rch[YP], rcl[YP];
  const float* Bq = B + q_off;
  auto gload = [&](int nu0) {
    const int nb = nu0 + 8 * kg;
#pragma unroll
    for (int j = 0; j < 8; ++j) rx[j] = *reinterpret_cast<const fl2*>(Bq + rowoff[nb + j]);
  };
  auto cload = [&](int nu0) {
#pragma unroll
    for (int jj = 0; jj < YP; ++jj) {
      const int e = tid + jj * NT;
      if (e < BI * 4) {
        const std::size_t off = static_cast<std::size_t>(i0 + (e >> 2)) * ldct + nu0 + (e & 3) * 8;
        rch[jj] = *reinterpret_cast<const bf8*>(Cth + off);
        rcl[jj] = *reinterpret_cast<const bf8*>(Ctl + off);
      }
    }
  };
  auto cstore = [&](int buf) {
#pragma unroll
    for (int jj = 0; jj < YP; ++jj) {
      const int e = tid + jj * NT;
      if (e < BI * 4) {
        const int yc = lds_chunk(e >> 2, e & 3) * 8;
        *reinterpret_cast<bf8*>(&Yh[buf][e >> 2][yc]) = rch[jj];
        *reinterpret_cast<bf8*>(&Yl[buf][e >> 2][yc]) = rcl[jj];
      }
    }
  };
  __syncthreads();  // rowoff and drow
  // Waves past ldq (only in the last Q tile) skip the gathers and MFMAs; they
  // still stage their share of C^T and meet every barrier.
  cload(0);
  if (q_ok) gload(0);
  cstore(0);
  __syncthreads();
  int cur = 0;
  for (int nu0 = 0; nu0 < n; nu0 += BK) {
    // This step's A fragments (waits for its gathers).  Lanes past ldq and
    // nu past n (only in the last step) contribute zeros.
    const int nb = nu0 + 8 * kg;
    float v0[8], v1[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      v0[j] = rx[j].x;
      v1[j] = rx[j].y;
    }
    if (!q_ok || nu0 + BK > n) {
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (!q_ok || nb + j >= n) v0[j] = v1[j] = 0.0f;
    }
    if (with_rho) {
      // rho partials: FP32 over the step's eight nu, FP64 across steps.
      float s0 = 0.0f, s1 = 0.0f;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const float d = nb + j < n ? drow[nb + j] : 0.0f;
        s0 = fmaf(d, v0[j], s0);
        s1 = fmaf(d, v1[j], s1);
      }
      racc[0] += static_cast<double>(s0);
      racc[1] += static_cast<double>(s1);
    }
    bf8 ah[XF], al[XF];
    split8_bf16(v0, ah[0], al[0]);
    split8_bf16(v1, ah[1], al[1]);
    const bool more = nu0 + BK < n;
    if (more) {
      cload(nu0 + BK);
      if (q_ok) gload(nu0 + BK);
    }
    const int kq = lds_chunk(lane & 15, kg) * 8;  // fragment rows start at multiples of 16
    if (q_ok) {
#pragma unroll
      for (int y = 0; y < YF; ++y) {
        const int row = y * 16 + (lane & 15);
        const bf8 bh = *reinterpret_cast<const bf8*>(&Yh[cur][row][kq]);
        const bf8 bl = *reinterpret_cast<const bf8*>(&Yl[cur][row][kq]);
#pragma unroll
        for (int x = 0; x < XF; ++x) acc[x][y] = mfma(ah[x], bh, acc[x][y]);
        if constexpr (TERMS == 3) {
#pragma unroll
          for (int x = 0; x < XF; ++x) acc[x][y] = mfma(ah[x], bl, acc[x][y]);
#pragma unroll
          for (int x = 0; x < XF; ++x) acc[x][y] = mfma(al[x], bh, acc[x][y]);
        }
      }
    }
    // The other buffer was last read before the previous barrier.
    if (more) cstore(cur ^ 1);
    __syncthreads();
    cur ^= 1;
  }
  if (with_rho) {
    // Sum the four k groups (lanes l, l ^ 16, l ^ 32, l ^ 48) of each row.
#pragma unroll
    for (int x = 0; x < XF; ++x) {
      double t = racc[x];
      t += __shfl_xor(t, 16);
      t += __shfl_xor(t, 32);
      if (kg == 0 && q_ok) rho_part[static_cast<std::size_t>(mu) * ldq + q_off + x] = t;
    }
  }
  // Fragment x row m holds Q = qw + 2 m + x; a lane's accumulator element v
  // is row m = 4 (l >> 4) + v of column i = l & 15.
  if (!q_ok) return;
#pragma unroll
  for (int y = 0; y < YF; ++y) {
    const int i = i0 + y * 16 + (lane & 15);
    if (i >= r) continue;
    const std::size_t off = static_cast<std::size_t>(mu) * us + static_cast<std::size_t>(i) * ldq + qw + 8 * (lane >> 4);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      split_bf16(acc[0][y][v], Uh[off + 2 * v], Ul[off + 2 * v]);
      split_bf16(acc[1][y][v], Uh[off + 2 * v + 1], Ul[off + 2 * v + 1]);
    }
  }
}

// ---------------------------------------------------------------------------
// K, step 2: lower-triangular split-K SYRK over W (n x kd, K-contiguous rows).
// Block: 256 threads = 2x2 waves, tile 64 x 64, partial results per split.
// ---------------------------------------------------------------------------


// FP64 / FP32 SYRK with the same block shape as the BF16X3 one (8 waves, tile
// 256 x 128, 64 x 64 per wave as 4 x 4 MFMA 16x16x4 tiles).  LDS rows keep W's
// K-contiguous order with one element of padding, so the 16 rows a wave reads
// at one k fall on distinct banks.
template <class T>
__global__ __launch_bounds__(512) void exchange_syrk_kernel(const T* __restrict__ W, std::size_t kd, int n,
                                                           int ntiles, int ncb, int nwork, std::size_t kbegin,
                                                           std::size_t kend, std::size_t kchunk, double sign,
                                                           T* __restrict__ partial);

// BF16X3 SYRK over the pre-split hi/lo planes written by occupied_u.
// Block: 8 waves (4 x 2), tile 256 (mu) x 128 (lambda); each wave owns 64 x 64
// as 4 x 4 MFMA 16x16x32 tiles, so one LDS k-step of 32 feeds 48 MFMAs per
// wave.  Row block rb covers min(2 (rb + 1), ncb) column blocks (the lower
// triangle).  Block ids are remapped so the blocks resident on one XCD
// (hardware dispatch is round-robin over the 8 XCDs) work on the same k-range
// of W and share that XCD's L2.
constexpr int kSyrkTM = 256, kSyrkTN = 128;  // FP64 / FP32 tiles
constexpr int kSyrkBfT = 256;                 // BF16X3 tiles (square)

// Lower-triangle tile t -> (row block, column block); row block rb holds
// min(ratio (rb + 1), ncb) column blocks, ratio = TM / TN.
__device__ __forceinline__ void syrk_tile_of(int t, int ncb, int ratio, int& rb, int& cb) {
  int base = 0;
  rb = 0;
  for (;;) {
    const int cnt = min(ratio * (rb + 1), ncb);
    if (t < base + cnt) {
      cb = t - base;
      return;
    }
    base += cnt;
    ++rb;
  }
}

// BF16X3 SYRK: 256 x 256 tiles, 8 waves as 2 (rows) x 4 (columns), each wave
// 128 x 64 = 8 x 4 MFMA 16x16x32 tiles.  Per k-step of 32 a wave reads 12
// fragment pairs from LDS for 96 MFMAs.  The next k-step is staged in
// registers while the current one runs (one LDS buffer, two barriers).
// CROSS: the column panels come from V instead of W and the tiles cover the
// whole n x n grid (ntiles = ncb^2): partial slices of W V^T, every element
// written.
template <bool CROSS, int TERMS>
__global__ __launch_bounds__(512) void exchange_syrk_bf16x3_kernel(const __bf16* __restrict__ Wh,
                                                                  const __bf16* __restrict__ Wl, std::size_t kd,
                                                                  int n, int ntiles, int ncb, int nwork,
                                                                  std::size_t kbegin, std::size_t kend,
                                                                  std::size_t kchunk, double sign,
                                                                  float* __restrict__ partial,
                                                                  const __bf16* __restrict__ Vh,
                                                                  const __bf16* __restrict__ Vl) {
  constexpr int T = kSyrkBfT, BK = 32, NT = 512, XF = 8, YF = 4;
  // Rows of 64 bytes with swizzled chunks (lds_chunk).
  __shared__ __bf16 Ah[T][BK], Al[T][BK];
  __shared__ __bf16 Bh[T][BK], Bl[T][BK];
  const int per = gridDim.x / 8;
  const int work = (blockIdx.x % 8) * per + blockIdx.x / 8;
  if (work >= nwork) return;
  const int split = work / ntiles, tile = work % ntiles;
  constexpr bool cross = CROSS;
  int rb, cb;
  if constexpr (CROSS) {
    rb = tile / ncb;
    cb = tile % ncb;
  } else {
    syrk_tile_of(tile, ncb, 1, rb, cb);
  }
  const int m0 = rb * T, l0 = cb * T;
  // A diagonal tile multiplies a panel by itself: one gather, one LDS copy.
  const bool diag = !cross && m0 == l0;
  const __bf16* __restrict__ Ch = cross ? Vh : Wh;  // column-panel operand
  const __bf16* __restrict__ Cl = cross ? Vl : Wl;
  const std::size_t k0 = kbegin + static_cast<std::size_t>(split) * kchunk;
  const std::size_t k1 = k0 + kchunk < kend ? k0 + kchunk : kend;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 2, wc = wave & 3;
  f4 acc[XF][YF];
#pragma unroll
  for (int i = 0; i < XF; ++i)
#pragma unroll
    for (int j = 0; j < YF; ++j) acc[i][j] = f4{};
  const bf8 zero = bf8{};
  constexpr int CH = T * 4 / NT;  // 16-byte chunks per thread, plane and operand
  bf8 ra[CH][2], rbv[CH][2];
  auto gload = [&](std::size_t kb) {
#pragma unroll
    for (int j = 0; j < CH; ++j) {
      const int c = tid + j * NT, row = c >> 2, kc = (c & 3) << 3;
      const int ga = m0 + row, gb = l0 + row;
      ra[j][0] = ra[j][1] = rbv[j][0] = rbv[j][1] = zero;
      if (ga < n) {
        const std::size_t off = static_cast<std::size_t>(ga) * kd + kb + kc;
        ra[j][0] = *reinterpret_cast<const bf8*>(Wh + off);
        ra[j][1] = *reinterpret_cast<const bf8*>(Wl + off);
      }
      if (!diag && gb < n) {
        const std::size_t off = static_cast<std::size_t>(gb) * kd + kb + kc;
        rbv[j][0] = *reinterpret_cast<const bf8*>(Ch + off);
        rbv[j][1] = *reinterpret_cast<const bf8*>(Cl + off);
      }
    }
  };
  auto sstore = [&]() {
#pragma unroll
    for (int j = 0; j < CH; ++j) {
      const int c = tid + j * NT, row = c >> 2, kc = lds_chunk(row, c & 3) << 3;
      *reinterpret_cast<bf8*>(&Ah[row][kc]) = ra[j][0];
      *reinterpret_cast<bf8*>(&Al[row][kc]) = ra[j][1];
      if (!diag) {
        *reinterpret_cast<bf8*>(&Bh[row][kc]) = rbv[j][0];
        *reinterpret_cast<bf8*>(&Bl[row][kc]) = rbv[j][1];
      }
    }
  };
  gload(k0);
  sstore();
  __syncthreads();
  for (std::size_t kb = k0; kb < k1; kb += BK) {
    const bool more = kb + BK < k1;
    if (more) gload(kb + BK);
    const int kq = lds_chunk(lane & 15, lane >> 4) * 8;  // fragment rows start at multiples of 16
    bf8 bh[YF], bl[YF];
#pragma unroll
    for (int j = 0; j < YF; ++j) {
      const int row = wc * 16 * YF + j * 16 + (lane & 15);
      bh[j] = *reinterpret_cast<const bf8*>(diag ? &Ah[row][kq] : &Bh[row][kq]);
      bl[j] = *reinterpret_cast<const bf8*>(diag ? &Al[row][kq] : &Bl[row][kq]);
    }
#pragma unroll
    for (int i = 0; i < XF; ++i) {
      const bf8 ah = *reinterpret_cast<const bf8*>(&Ah[wr * 16 * XF + i * 16 + (lane & 15)][kq]);
      const bf8 al = *reinterpret_cast<const bf8*>(&Al[wr * 16 * XF + i * 16 + (lane & 15)][kq]);
      // Independent accumulators between dependent MFMAs keep the pipe full.
#pragma unroll
      for (int j = 0; j < YF; ++j) acc[i][j] = mfma(ah, bh[j], acc[i][j]);
      if constexpr (TERMS == 3) {
#pragma unroll
        for (int j = 0; j < YF; ++j) acc[i][j] = mfma(ah, bl[j], acc[i][j]);
#pragma unroll
        for (int j = 0; j < YF; ++j) acc[i][j] = mfma(al, bh[j], acc[i][j]);
      }
    }
    __syncthreads();
    if (more) {
      sstore();
      __syncthreads();
    }
  }
  // FP32 partials hold the FP32 accumulators exactly (sign is +-1).
  float* out = partial + static_cast<std::size_t>(split) * n * n;
  const float fsign = static_cast<float>(sign);
#pragma unroll
  for (int i = 0; i < XF; ++i)
#pragma unroll
    for (int j = 0; j < YF; ++j) {
      const int lam = l0 + wc * 16 * YF + j * 16 + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int mu = m0 + wr * 16 * XF + i * 16 + 4 * (lane >> 4) + v;
        if (mu < n && lam < n && (cross || lam <= mu)) out[static_cast<std::size_t>(mu) * n + lam] = fsign * acc[i][j][v];
      }
    }
}

template <class T>
__global__ __launch_bounds__(512) void exchange_syrk_kernel(const T* __restrict__ W, std::size_t kd, int n,
                                                           int ntiles, int ncb, int nwork, std::size_t kbegin,
                                                           std::size_t kend, std::size_t kchunk, double sign,
                                                           T* __restrict__ partial) {
  using Acc = typename AccOf<T>::type;
  constexpr int TM = kSyrkTM, TN = kSyrkTN, BK = 16, PAD = 1, NT = 512, V = 16 / sizeof(T);
  __shared__ T As[2][TM][BK + PAD];
  __shared__ T Bs[2][TN][BK + PAD];
  const int per = gridDim.x / 8;
  const int work = (blockIdx.x % 8) * per + blockIdx.x / 8;
  if (work >= nwork) return;
  const int split = work / ntiles, tile = work % ntiles;
  int rb, cb;
  syrk_tile_of(tile, ncb, TM / TN, rb, cb);
  const int m0 = rb * TM, l0 = cb * TN;
  const std::size_t k0 = kbegin + static_cast<std::size_t>(split) * kchunk;
  const std::size_t k1 = k0 + kchunk < kend ? k0 + kchunk : kend;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  Acc acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = Acc{};
  constexpr int CH = BK / V;                  // 16-byte chunks per row
  constexpr int AC = TM * CH / NT, BC = TN * CH / NT;
  T ra[AC][V], rbv[BC][V];
  // Next k-step staged in registers, LDS double-buffered: one barrier per step.
  auto gload = [&](std::size_t kb) {
#pragma unroll
    for (int j = 0; j < AC; ++j) {
      const int c = tid + j * NT, row = c / CH, kc = (c % CH) * V, grow = m0 + row;
#pragma unroll
      for (int e = 0; e < V; ++e) ra[j][e] = grow < n ? W[static_cast<std::size_t>(grow) * kd + kb + kc + e] : T(0);
    }
#pragma unroll
    for (int j = 0; j < BC; ++j) {
      const int c = tid + j * NT, row = c / CH, kc = (c % CH) * V, grow = l0 + row;
#pragma unroll
      for (int e = 0; e < V; ++e) rbv[j][e] = grow < n ? W[static_cast<std::size_t>(grow) * kd + kb + kc + e] : T(0);
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int j = 0; j < AC; ++j) {
      const int c = tid + j * NT, row = c / CH, kc = (c % CH) * V;
#pragma unroll
      for (int e = 0; e < V; ++e) As[buf][row][kc + e] = ra[j][e];
    }
#pragma unroll
    for (int j = 0; j < BC; ++j) {
      const int c = tid + j * NT, row = c / CH, kc = (c % CH) * V;
#pragma unroll
      for (int e = 0; e < V; ++e) Bs[buf][row][kc + e] = rbv[j][e];
    }
  };
  gload(k0);
  sstore(0);
  __syncthreads();
  int cur = 0;
  for (std::size_t kb = k0; kb < k1; kb += BK) {
    const bool more = kb + BK < k1;
    if (more) gload(kb + BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int k = kk + (lane >> 4);
      T b[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[cur][wc * 64 + j * 16 + (lane & 15)][k];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const T a = As[cur][wr * 64 + i * 16 + (lane & 15)][k];
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = mfma(a, b[j], acc[i][j]);
      }
    }
    if (more) sstore(cur ^ 1);
    __syncthreads();
    cur ^= 1;
  }
  // Partials in the accumulator's own type hold it exactly (sign is +-1).
  T* out = partial + static_cast<std::size_t>(split) * n * n;
  const T tsign = static_cast<T>(sign);
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int lam = l0 + wc * 64 + j * 16 + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int mu = m0 + wr * 64 + i * 16 + acc_row<T>(lane, v);
        if (mu < n && lam < n && lam <= mu) out[static_cast<std::size_t>(mu) * n + lam] = tsign * acc[i][j][v];
      }
    }
}

// K = sum of the split partials (FP64 sum in slice order), mirrored.
template <class P>
__global__ void reduce_exchange_kernel(const P* __restrict__ partial, int slices, int n, double* __restrict__ K) {
  const std::size_t e = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= static_cast<std::size_t>(n) * n) return;
  const int mu = static_cast<int>(e / n), lam = static_cast<int>(e % n);
  if (lam > mu) return;
  double s = 0.0;
  for (int k = 0; k < slices; ++k) s += static_cast<double>(partial[static_cast<std::size_t>(k) * n * n + e]);
  K[static_cast<std::size_t>(mu) * n + lam] = s;
  K[static_cast<std::size_t>(lam) * n + mu] = s;
}

// K = Kref + W + W^T with W the sum of the cross partial slices (FP64 sum in
// slice order); K is symmetric by construction.
__global__ void reduce_cross_kernel(const float* __restrict__ partial, int slices, int n,
                                    const double* __restrict__ Kref, double* __restrict__ K) {
  const std::size_t e = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= static_cast<std::size_t>(n) * n) return;
  const int mu = static_cast<int>(e / n), lam = static_cast<int>(e % n);
  if (lam > mu) return;
  const std::size_t et = static_cast<std::size_t>(lam) * n + mu;
  double s = 0.0, t = 0.0;
  for (int k = 0; k < slices; ++k) {
    s += static_cast<double>(partial[static_cast<std::size_t>(k) * n * n + e]);
    t += static_cast<double>(partial[static_cast<std::size_t>(k) * n * n + et]);
  }
  const double v = Kref[e] + (s + t);
  K[e] = v;
  K[et] = v;
}

__global__ void convert_kernel(const double* __restrict__ in, float* __restrict__ out, std::size_t count) {
  const std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const std::size_t stride = static_cast<std::size_t>(gridDim.x) * blockDim.x;
  for (std::size_t k = i; k < count; k += stride) out[k] = static_cast<float>(in[k]);
}

// C (n x ldc, FP64) -> C^T split: hi/lo bf16, ldc rows x ldct columns (zero beyond n).
__global__ void split_transpose_kernel(const double* __restrict__ C, int n, int ldc, int ldct,
                                       __bf16* __restrict__ hi, __bf16* __restrict__ lo) {
  const std::size_t e = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= static_cast<std::size_t>(ldc) * ldct) return;
  const int i = static_cast<int>(e / ldct), nu = static_cast<int>(e % ldct);
  const float x = nu < n ? static_cast<float>(C[static_cast<std::size_t>(nu) * ldc + i]) : 0.0f;
  __bf16 h, l;
  split_bf16(x, h, l);
  hi[e] = h;
  lo[e] = l;
}

template <int NIW>
void launch_u(Precision prec, const void* B, const OccupiedOperands& ops, int ldq, void* U, const double* D,
              double* rho_part, hipStream_t s) {
  const std::size_t dyn = rho_part ? static_cast<std::size_t>(ops.n) * sizeof(double) : 0;
  if (dyn > 32 * 1024) throw NotImplemented("fused rho supports nbf <= 4096");
  if (prec == Precision::FP64) {
    constexpr int WQ = 4;
    const int nqb = (ldq + 32 * WQ - 1) / (32 * WQ);
    dim3 grid(static_cast<unsigned>(ops.n) * nqb, ops.ldc / (32 * NIW));
    hipLaunchKernelGGL((occupied_u_kernel<double, NIW, WQ>), grid, dim3(128 * WQ), dyn, s,
                       static_cast<const double*>(B), ops.c64, static_cast<double*>(U), ops.ustride, ops.n, ldq,
                       ops.ldc, ops.r, nqb, D, rho_part);
  } else if (prec == Precision::FP32) {
    constexpr int WQ = 8;
    const int nqb = (ldq + 32 * WQ - 1) / (32 * WQ);
    dim3 grid(static_cast<unsigned>(ops.n) * nqb, ops.ldc / (32 * NIW));
    hipLaunchKernelGGL((occupied_u_kernel<float, NIW, WQ>), grid, dim3(128 * WQ), dyn, s, static_cast<const float*>(B),
                       ops.c32, static_cast<float*>(U), ops.ustride, ops.n, ldq, ops.ldc, ops.r, nqb, D, rho_part);
  } else if constexpr (NIW <= 10) {
    const int nqb = (ldq + 255) / 256;
    dim3 grid(static_cast<unsigned>(ops.n) * nqb, ops.ldc / (32 * NIW));
    // Row-offset table (n + 32 entries) and the rho row as float.
    const std::size_t ddyn = static_cast<std::size_t>(ops.n + 32) * sizeof(std::size_t) +
                             (rho_part ? static_cast<std::size_t>(ops.n) * sizeof(float) : 0);
    auto kernel = ops.terms == 1 ? occupied_u_bf16x3_direct_kernel<NIW, 1> : occupied_u_bf16x3_direct_kernel<NIW, 3>;
    hipLaunchKernelGGL(kernel, grid, dim3(512), ddyn, s, static_cast<const float*>(B),
                       static_cast<const __bf16*>(ops.ct_hi), static_cast<const __bf16*>(ops.ct_lo),
                       static_cast<__bf16*>(U), static_cast<__bf16*>(U) + static_cast<std::size_t>(ops.n) * ops.ustride,
                       ops.ustride, ops.n, ldq, ops.ldct, ops.r, D, rho_part);
  } else {
    if (ops.terms != 3) throw std::logic_error("single-term BF16 products need i tiles of at most 320 columns");
    const int nqb = (ldq + kUTileQ - 1) / kUTileQ;
    dim3 grid(static_cast<unsigned>(ops.n) * nqb, ops.ldc / (32 * NIW));
    // 8 x 2 waves: the 4 x 4 layout needs more than the 128 registers a
    // 16-wave block allows once NIW >= 8.
    hipLaunchKernelGGL((occupied_u_bf16x3_kernel<NIW, 8, (NIW <= 4 ? 2 : 1)>), grid, dim3(1024), dyn, s, static_cast<const float*>(B),
                       static_cast<const __bf16*>(ops.ct_hi), static_cast<const __bf16*>(ops.ct_lo),
                       static_cast<__bf16*>(U), static_cast<__bf16*>(U) + static_cast<std::size_t>(ops.n) * ops.ustride,
                       ops.ustride, ops.n, ldq, ops.ldct, ops.r, nqb, D, rho_part);
  }
  VQC_LAUNCH("occupied_u_kernel");
}


}  // namespace

int occupied_tile(int r) {
  // i-tiles of 16 per wave, two waves along i: one block covers 32 * NIW
  // columns, all of r up to 320.
  return std::max(1, std::min(10, (r + 31) / 32));
}

void pack_density(const double* D, int n, double* dpacked, hipStream_t s) {
  const std::size_t np = pair_count(n);
  hipLaunchKernelGGL(pack_density_kernel, dim3((np + 255) / 256), dim3(256), 0, s, D, n, dpacked);
  VQC_LAUNCH("pack_density_kernel");
}

namespace {
int rho_vector(Precision prec) { return prec == Precision::FP64 ? 2 : 4; }
int rho_chunks(std::size_t npair, int ldq, Precision prec) {
  const int cols = (ldq + kRhoThreads * rho_vector(prec) - 1) / (kRhoThreads * rho_vector(prec));
  const int want = std::max(1, 2048 / cols);
  return static_cast<int>(std::min<std::size_t>(want, std::max<std::size_t>(1, npair / 64)));
}
}  // namespace

std::size_t coulomb_partial_size(std::size_t npair, int ldq) {
  return static_cast<std::size_t>(std::max(rho_chunks(npair, ldq, Precision::FP64),
                                           rho_chunks(npair, ldq, Precision::FP32))) *
         ldq;
}

void coulomb_rho(Precision prec, const void* B, const double* dpacked, std::size_t npair, int ldq, double* rho,
                 double* partial, hipStream_t s) {
  const int v = rho_vector(prec);
  const int chunks = rho_chunks(npair, ldq, prec);
  const std::size_t chunk = (npair + chunks - 1) / chunks;
  dim3 grid((ldq + kRhoThreads * v - 1) / (kRhoThreads * v), chunks);
  if (prec == Precision::FP64)
    hipLaunchKernelGGL((coulomb_rho_kernel<double, 2>), grid, dim3(kRhoThreads), 0, s,
                       static_cast<const double*>(B), dpacked, npair, ldq, chunk, partial);
  else
    hipLaunchKernelGGL((coulomb_rho_kernel<float, 4>), grid, dim3(kRhoThreads), 0, s,
                       static_cast<const float*>(B), dpacked, npair, ldq, chunk, partial);
  VQC_LAUNCH("coulomb_rho_kernel");
  hipLaunchKernelGGL(sum_rows_kernel, dim3((ldq + kSumCols - 1) / kSumCols), dim3(kSumCols * kSumGroups), 0, s, partial,
                     chunks, ldq, rho);
  VQC_LAUNCH("sum_rows_kernel");
}

void rho_from_partials(const double* rho_part, int rows, int ldq, double* rho, hipStream_t s) {
  hipLaunchKernelGGL(sum_rows_kernel, dim3((ldq + kSumCols - 1) / kSumCols), dim3(kSumCols * kSumGroups), 0, s,
                     rho_part, rows, ldq, rho);
  VQC_LAUNCH("sum_rows_kernel");
}

void coulomb_matrix(Precision prec, const void* B, const double* rho, std::size_t npair, int ldq, int n, double* J,
                    hipStream_t s) {
  const std::size_t lds = static_cast<std::size_t>(ldq) * sizeof(double);
  if (lds > 160 * 1024) throw NotImplemented("auxiliary dimension exceeds the LDS-resident rho limit");
  // Up to two resident blocks per CU on 256 CUs (one when rho exceeds 80 KB);
  // waves stride over the packed rows.
  const std::size_t rows_per_block = kJThreads / 64;
  const unsigned blocks = static_cast<unsigned>(std::min<std::size_t>((npair + rows_per_block - 1) / rows_per_block, 512));
  if (prec == Precision::FP64)
    hipLaunchKernelGGL(coulomb_matrix_kernel<double>, dim3(blocks), dim3(kJThreads), lds, s,
                       static_cast<const double*>(B), rho, npair, ldq, n, J);
  else
    hipLaunchKernelGGL(coulomb_matrix_kernel<float>, dim3(blocks), dim3(kJThreads), lds, s,
                       static_cast<const float*>(B), rho, npair, ldq, n, J);
  VQC_LAUNCH("coulomb_matrix_kernel");
}

void prepare_occupied(Precision prec, OccupiedOperands& ops, hipStream_t s) {
  const std::size_t count = static_cast<std::size_t>(ops.n) * ops.ldc;
  if (prec == Precision::FP32) {
    hipLaunchKernelGGL(convert_kernel, dim3(std::min<std::size_t>((count + 255) / 256, 4096)), dim3(256), 0, s,
                       ops.c64, ops.c32, count);
    VQC_LAUNCH("convert_kernel");
  } else if (prec == Precision::BF16X3) {
    const std::size_t e = static_cast<std::size_t>(ops.ldc) * ops.ldct;
    hipLaunchKernelGGL(split_transpose_kernel, dim3((e + 255) / 256), dim3(256), 0, s, ops.c64, ops.n, ops.ldc,
                       ops.ldct, static_cast<__bf16*>(ops.ct_hi), static_cast<__bf16*>(ops.ct_lo));
    VQC_LAUNCH("split_transpose_kernel");
  }
}

void occupied_u(Precision prec, const void* B, const OccupiedOperands& ops, int ldq, void* U, const double* D,
                double* rho_part, hipStream_t s) {
  switch (ops.ni) {
    case 1: launch_u<1>(prec, B, ops, ldq, U, D, rho_part, s); break;
    case 2: launch_u<2>(prec, B, ops, ldq, U, D, rho_part, s); break;
    case 3: launch_u<3>(prec, B, ops, ldq, U, D, rho_part, s); break;
    case 4: launch_u<4>(prec, B, ops, ldq, U, D, rho_part, s); break;
    case 5: launch_u<5>(prec, B, ops, ldq, U, D, rho_part, s); break;
    case 6: launch_u<6>(prec, B, ops, ldq, U, D, rho_part, s); break;
    case 7: launch_u<7>(prec, B, ops, ldq, U, D, rho_part, s); break;
    case 8: launch_u<8>(prec, B, ops, ldq, U, D, rho_part, s); break;
    case 9: launch_u<9>(prec, B, ops, ldq, U, D, rho_part, s); break;
    case 10: launch_u<10>(prec, B, ops, ldq, U, D, rho_part, s); break;
    default: throw std::logic_error("occupied tile width out of range");
  }
}

namespace {
// Split-K slices per launch: at most 256 MB of FP64 partial slices (and 128).
int syrk_split_cap(int n) {
  const std::size_t slice = static_cast<std::size_t>(n) * n * sizeof(double);
  return static_cast<int>(std::max<std::size_t>(1, std::min<std::size_t>(128, (256ull << 20) / slice)));
}
int syrk_tiles(int n, int tm, int tn, int& ncb) {
  ncb = (n + tn - 1) / tn;
  const int nrb = (n + tm - 1) / tm;
  int t = 0;
  for (int rb = 0; rb < nrb; ++rb) t += std::min((tm / tn) * (rb + 1), ncb);
  return t;
}
// Equal-size work items fill the GPU in whole waves of `slots` resident
// blocks: take the fewest splits whose tiles x splits fill at least 95% of the
// last wave, else the best fill under the cap.
int syrk_splits(int tiles, int cap, int slots) {
  const int kSlots = slots;
  int best = 1;
  double best_fill = 0.0;
  for (int sp = 1; sp <= cap; ++sp) {
    const int w = tiles * sp;
    const double fill = static_cast<double>(w) / (static_cast<double>((w + kSlots - 1) / kSlots) * kSlots);
    if (w >= kSlots && fill >= 0.95) return sp;
    if (fill > best_fill + 1e-12) {
      best = sp;
      best_fill = fill;
    }
  }
  return best;
}
}  // namespace

std::size_t exchange_partial_size(int n) {
  return static_cast<std::size_t>(syrk_split_cap(n) + 1) * 2 * static_cast<std::size_t>(n) * n;
}

std::size_t u_row_stride(Precision prec, int r, int ldq) {
  const std::size_t dense = static_cast<std::size_t>(r) * ldq;
  if (prec == Precision::FP64) return dense;
  return dense + (prec == Precision::FP32 ? 32 : 64);
}

namespace {
// Slices per run: both runs (positive and negative columns) together must
// fit the partial buffer (n^2 doubles per FP64 slice; FP32 slices use half).
int syrk_cap(int n, std::size_t partial_capacity) {
  const std::size_t slice = static_cast<std::size_t>(n) * n;
  return partial_capacity == 0
             ? syrk_split_cap(n)
             : static_cast<int>(std::min<std::size_t>(syrk_split_cap(n), partial_capacity / (2 * slice)));
}
int syrk_tile_count(Precision prec, int n, int& ncb) {
  return prec == Precision::BF16X3 ? syrk_tiles(n, kSyrkBfT, kSyrkBfT, ncb) : syrk_tiles(n, kSyrkTM, kSyrkTN, ncb);
}
// Resident 512-thread blocks on 256 CUs: one per CU for BF16X3 (64 KB LDS)
// and FP64 (102 KB), two for FP32 (51 KB).
int syrk_slots(Precision prec) { return prec == Precision::FP32 ? 512 : 256; }
}  // namespace

int exchange_max_splits(int n, std::size_t partial_capacity) { return std::max(1, syrk_cap(n, partial_capacity)); }

int exchange_default_splits(Precision prec, int n, std::size_t partial_capacity) {
  const int cap = syrk_cap(n, partial_capacity);
  if (cap < 1) throw std::length_error("exchange partial buffer holds no slice");
  int ncb = 0;
  return syrk_splits(syrk_tile_count(prec, n, ncb), cap, syrk_slots(prec));
}

void exchange_syrk(Precision prec, const void* U, std::size_t ustride, int n, int r, int r_pos, int ldq, double* K,
                   double* partial, hipStream_t s, std::size_t partial_capacity, int split_override, int terms) {
  const std::size_t kd = static_cast<std::size_t>(r) * ldq;
  if (ustride < kd) throw std::invalid_argument("U row stride shorter than r x ldq");
  const int cap = syrk_cap(n, partial_capacity);
  if (cap < 1) throw std::length_error("exchange partial buffer holds no slice");
  int slices = 0;
  auto run = [&](std::size_t kb, std::size_t ke, double sign) {
    if (ke <= kb) return;
    const std::size_t len = ke - kb;
    // FP64 partials for the FP64 build, FP32 (the accumulators' type) otherwise.
    const std::size_t at = static_cast<std::size_t>(slices) * n * n;
    double* out64 = partial + at;
    float* out32 = reinterpret_cast<float*>(partial) + at;
    int ncb = 0;
    const int tiles = syrk_tile_count(prec, n, ncb);
    const std::size_t bk = prec == Precision::BF16X3 ? 32 : 16;
    int splits = split_override > 0 ? std::min(split_override, cap) : syrk_splits(tiles, cap, syrk_slots(prec));
    const std::size_t chunk = round_up((len + splits - 1) / splits, bk);
    splits = static_cast<int>((len + chunk - 1) / chunk);
    const int nwork = tiles * splits;
    const unsigned grid = static_cast<unsigned>(round_up(nwork, 8));
    if (prec == Precision::BF16X3) {
      const __bf16* wh = static_cast<const __bf16*>(U);
      const __bf16* wl = wh + static_cast<std::size_t>(n) * ustride;
      auto kernel = terms == 1 ? exchange_syrk_bf16x3_kernel<false, 1> : exchange_syrk_bf16x3_kernel<false, 3>;
      hipLaunchKernelGGL(kernel, dim3(grid), dim3(512), 0, s, wh, wl, ustride, n, tiles, ncb, nwork, kb, ke, chunk, sign,
                         out32, wh, wl);
    } else if (prec == Precision::FP64) {
      hipLaunchKernelGGL(exchange_syrk_kernel<double>, dim3(grid), dim3(512), 0, s, static_cast<const double*>(U),
                         ustride, n, tiles, ncb, nwork, kb, ke, chunk, sign, out64);
    } else {
      hipLaunchKernelGGL(exchange_syrk_kernel<float>, dim3(grid), dim3(512), 0, s, static_cast<const float*>(U),
                         ustride, n, tiles, ncb, nwork, kb, ke, chunk, sign, out32);
    }
    VQC_LAUNCH("exchange_syrk_kernel");
    slices += splits;
  };
  run(0, static_cast<std::size_t>(r_pos) * ldq, 1.0);
  run(static_cast<std::size_t>(r_pos) * ldq, kd, -1.0);
  if (slices == 0) {
    VQC_HIP(hipMemsetAsync(K, 0, static_cast<std::size_t>(n) * n * sizeof(double), s));
    return;
  }
  const std::size_t e = static_cast<std::size_t>(n) * n;
  if (prec == Precision::FP64)
    hipLaunchKernelGGL(reduce_exchange_kernel<double>, dim3((e + 255) / 256), dim3(256), 0, s, partial, slices, n, K);
  else
    hipLaunchKernelGGL(reduce_exchange_kernel<float>, dim3((e + 255) / 256), dim3(256), 0, s,
                       reinterpret_cast<const float*>(partial), slices, n, K);
  VQC_LAUNCH("reduce_exchange_kernel");
}

void exchange_cross_bf16x3(const void* X, const void* Y, std::size_t ustride, int n, int r, int ldq,
                           const double* Kref, double* K, double* partial, hipStream_t s) {
  const std::size_t kd = static_cast<std::size_t>(r) * ldq;
  if (ustride < kd) throw std::invalid_argument("U row stride shorter than r x ldq");
  const int cap = syrk_cap(n, 0);
  const int ncb = (n + kSyrkBfT - 1) / kSyrkBfT;
  const int tiles = ncb * ncb;
  int splits = syrk_splits(tiles, cap, syrk_slots(Precision::BF16X3));
  const std::size_t chunk = round_up((kd + splits - 1) / splits, 32);
  splits = static_cast<int>((kd + chunk - 1) / chunk);
  if (splits > 2 * cap) throw std::length_error("cross partial slices exceed the partial buffer");
  const int nwork = tiles * splits;
  const unsigned grid = static_cast<unsigned>(round_up(nwork, 8));
  const __bf16* xh = static_cast<const __bf16*>(X);
  const __bf16* xl = xh + static_cast<std::size_t>(n) * ustride;
  const __bf16* yh = static_cast<const __bf16*>(Y);
  const __bf16* yl = yh + static_cast<std::size_t>(n) * ustride;
  float* out32 = reinterpret_cast<float*>(partial);
  hipLaunchKernelGGL((exchange_syrk_bf16x3_kernel<true, 3>), dim3(grid), dim3(512), 0, s, xh, xl, ustride, n, tiles, ncb, nwork,
                     std::size_t{0}, kd, chunk, 1.0, out32, yh, yl);
  VQC_LAUNCH("exchange_syrk_bf16x3_kernel (cross)");
  const std::size_t e = static_cast<std::size_t>(n) * n;
  hipLaunchKernelGGL(reduce_cross_kernel, dim3((e + 255) / 256), dim3(256), 0, s, out32, splits, n, Kref, K);
  VQC_LAUNCH("reduce_cross_kernel");
}

namespace {
__global__ void occupied_factor_kernel(const double* __restrict__ C, int n, int nocc, int ldc,
                                       double* __restrict__ factor) {
  const std::size_t e = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= static_cast<std::size_t>(n) * ldc) return;
  const int nu = static_cast<int>(e / ldc), i = static_cast<int>(e % ldc);
  factor[e] = i < nocc ? 1.4142135623730951 * C[static_cast<std::size_t>(nu) * nocc + i] : 0.0;
}

// D[mu][nu] = sum_i F[mu][i] F[nu][i], one element per thread (fixed order).
__global__ void factor_product_kernel(const double* __restrict__ F, int n, int nocc, int ldc, double* __restrict__ D) {
  const std::size_t e = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= static_cast<std::size_t>(n) * n) return;
  const int mu = static_cast<int>(e / n), nu = static_cast<int>(e % n);
  const double* a = F + static_cast<std::size_t>(mu) * ldc;
  const double* b = F + static_cast<std::size_t>(nu) * ldc;
  double s = 0.0;
  for (int i = 0; i < nocc; ++i) s += a[i] * b[i];
  D[e] = s;
}
}  // namespace

void occupied_factor_and_density(const double* C, int n, int nocc, int ldc, double* factor, double* D,
                                 hipStream_t s) {
  const std::size_t ef = static_cast<std::size_t>(n) * ldc, ed = static_cast<std::size_t>(n) * n;
  hipLaunchKernelGGL(occupied_factor_kernel, dim3((ef + 255) / 256), dim3(256), 0, s, C, n, nocc, ldc, factor);
  VQC_LAUNCH("occupied_factor_kernel");
  hipLaunchKernelGGL(factor_product_kernel, dim3((ed + 255) / 256), dim3(256), 0, s, factor, n, nocc, ldc, D);
  VQC_LAUNCH("factor_product_kernel");
}

void convert_f64_to_f32(const double* in, float* out, std::size_t count, hipStream_t s) {
  hipLaunchKernelGGL(convert_kernel, dim3(std::min<std::size_t>((count + 255) / 256, 65536)), dim3(256), 0, s, in,
                     out, count);
  VQC_LAUNCH("convert_kernel");
}

}  // namespace vqc
