// Device kernels of the RHF SCF loop: Fock assembly, energy contractions,
// DIIS (Pulay, FDS - SDF residual; src/scf/solver/diis.cpp:17-46,
// src/solver/diis_coefficients.hpp) and orbital/factor bookkeeping.
// Reductions run in one 1024-thread block with a fixed summation order, so
// every result is bit-reproducible run to run.
#include "scf.hpp"

namespace vqc {
namespace {

constexpr int kReduceThreads = 1024;

__device__ double block_sum(double v, double* shared) {
  const int tid = threadIdx.x;
  shared[tid] = v;
  __syncthreads();
  for (int off = kReduceThreads / 2; off > 0; off >>= 1) {
    if (tid < off) shared[tid] += shared[tid + off];
    __syncthreads();
  }
  const double r = shared[0];
  __syncthreads();
  return r;
}

__global__ void assemble_fock_kernel(const double* __restrict__ H, const double* __restrict__ J,
                                     const double* __restrict__ K, double cj, double ck, std::size_t count,
                                     double* __restrict__ F) {
  const std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < count) F[i] = H[i] + cj * J[i] + ck * K[i];
}

__global__ __launch_bounds__(kReduceThreads) void dot_pair_kernel(const double* __restrict__ D,
                                                                  const double* __restrict__ A,
                                                                  const double* __restrict__ B,
                                                                  std::size_t count, double* __restrict__ out) {
  __shared__ double sh[kReduceThreads];
  double sa = 0.0, sb = 0.0;
  for (std::size_t i = threadIdx.x; i < count; i += kReduceThreads) {
    sa += D[i] * A[i];
    sb += D[i] * B[i];
  }
  sa = block_sum(sa, sh);
  sb = block_sum(sb, sh);
  if (threadIdx.x == 0) {
    out[0] = sa;
    out[1] = sb;
  }
}

__global__ __launch_bounds__(kReduceThreads) void rms_kernel(const double* __restrict__ a,
                                                             const double* __restrict__ b, std::size_t count,
                                                             double* __restrict__ out) {
  __shared__ double sh[kReduceThreads];
  double s = 0.0;
  for (std::size_t i = threadIdx.x; i < count; i += kReduceThreads) {
    const double d = a[i] - b[i];
    s += d * d;
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) out[0] = sqrt(s / static_cast<double>(count));
}

__global__ void antisymmetrise_kernel(const double* __restrict__ T, int n, double* __restrict__ R) {
  const std::size_t e = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= static_cast<std::size_t>(n) * n) return;
  const int i = static_cast<int>(e / n), j = static_cast<int>(e % n);
  R[e] = T[e] - T[static_cast<std::size_t>(j) * n + i];
}

__global__ __launch_bounds__(kReduceThreads) void diis_gram_kernel(const double* __restrict__ errors,
                                                                   std::size_t count, int row_slot,
                                                                   const int* __restrict__ slot_of, int m,
                                                                   int cap, double* __restrict__ gram) {
  __shared__ double sh[kReduceThreads];
  const int j = blockIdx.x;
  if (j >= m) return;
  const int col_slot = slot_of[j];
  const double* a = errors + static_cast<std::size_t>(row_slot) * count;
  const double* b = errors + static_cast<std::size_t>(col_slot) * count;
  double s = 0.0;
  for (std::size_t i = threadIdx.x; i < count; i += kReduceThreads) s += a[i] * b[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) {
    gram[row_slot * cap + col_slot] = s;
    gram[col_slot * cap + row_slot] = s;
  }
}

// Augmented Pulay system [G -1; -1 0][c; l] = [0; -1] by Gaussian elimination
// with partial pivoting (m <= 32), one wavefront: the system lives in LDS,
// lanes own rows for the elimination and columns for the swaps, and lane 0
// picks pivots and back-substitutes, so the result is deterministic and the
// kernel needs no private (scratch) memory.
__global__ __launch_bounds__(64) void diis_solve_kernel(const double* __restrict__ gram, int m,
                                                        const int* __restrict__ slot_of, int cap,
                                                        double* __restrict__ coef) {
  constexpr int kMax = 33;
  __shared__ double a[kMax][kMax + 1];
  __shared__ double x[kMax];
  __shared__ int piv_row, ok_flag;
  const int lane = threadIdx.x;
  const int d = m + 1;
  for (int e = lane; e < d * (d + 1); e += 64) {
    const int i = e / (d + 1), j = e % (d + 1);
    double v;
    if (i < m && j < m) v = gram[slot_of[i] * cap + slot_of[j]];
    else if (i < m) v = j == m ? -1.0 : 0.0;       // column m: -1, rhs: 0
    else v = j < m ? -1.0 : (j == m ? 0.0 : -1.0);  // last row: -1 ... 0 | -1
    a[i][j] = v;
  }
  if (lane == 0) ok_flag = 1;
  __syncthreads();
  double scale = 0.0;
  if (lane == 0)
    for (int i = 0; i < d; ++i)
      for (int j = 0; j < d; ++j) scale = fmax(scale, fabs(a[i][j]));
  for (int c = 0; c < d; ++c) {
    if (lane == 0) {
      int p = c;
      for (int r = c + 1; r < d; ++r)
        if (fabs(a[r][c]) > fabs(a[p][c])) p = r;
      piv_row = p;
      if (!(fabs(a[p][c]) > 1e-300) || !(fabs(a[p][c]) > scale * 1e-300)) ok_flag = 0;
    }
    __syncthreads();
    if (!ok_flag) break;
    const int p = piv_row;
    if (p != c && lane <= d) {
      const double t = a[c][lane];
      a[c][lane] = a[p][lane];
      a[p][lane] = t;
    }
    __syncthreads();
    if (lane > c && lane < d) {
      const double f = a[lane][c] / a[c][c];
      for (int j = c; j <= d; ++j) a[lane][j] -= f * a[c][j];
    }
    __syncthreads();
  }
  if (lane == 0) {
    bool ok = ok_flag != 0;
    if (ok) {
      for (int i = d - 1; i >= 0; --i) {
        double t = a[i][d];
        for (int j = i + 1; j < d; ++j) t -= a[i][j] * x[j];
        x[i] = t / a[i][i];
      }
      for (int i = 0; i < m; ++i)
        if (!isfinite(x[i])) ok = false;
    }
    for (int i = 0; i < m; ++i) coef[i] = ok ? x[i] : (i == m - 1 ? 1.0 : 0.0);
  }
}

__global__ void diis_extrapolate_kernel(const double* __restrict__ focks, std::size_t count, int m,
                                        const int* __restrict__ slot_of, const double* __restrict__ coef,
                                        double* __restrict__ F) {
  const std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  double s = 0.0;
  for (int j = 0; j < m; ++j) s += coef[j] * focks[static_cast<std::size_t>(slot_of[j]) * count + i];
  F[i] = s;
}

__global__ void scale_columns_kernel(const double* __restrict__ V, const double* __restrict__ w, int n,
                                     double* __restrict__ W) {
  const std::size_t e = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= static_cast<std::size_t>(n) * n) return;
  const int col = static_cast<int>(e / n);
  W[e] = V[e] / sqrt(w[col]);
}

__global__ void occupied_kernel(const double* __restrict__ C, int n, int nocc, int ldc, double scale,
                                double* __restrict__ Cocc) {
  const std::size_t e = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= static_cast<std::size_t>(n) * ldc) return;
  const int nu = static_cast<int>(e / ldc), i = static_cast<int>(e % ldc);
  Cocc[e] = i < nocc ? scale * C[nu + static_cast<std::size_t>(i) * n] : 0.0;
}

__global__ void factor_kernel(const double* __restrict__ V, const double* __restrict__ w,
                              const int* __restrict__ idx, int n, int r, int ldc, double* __restrict__ Cocc) {
  const std::size_t e = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= static_cast<std::size_t>(n) * ldc) return;
  const int nu = static_cast<int>(e / ldc), i = static_cast<int>(e % ldc);
  if (i >= r) {
    Cocc[e] = 0.0;
    return;
  }
  const int k = idx[i];
  Cocc[e] = sqrt(fabs(w[k])) * V[nu + static_cast<std::size_t>(k) * n];
}

__global__ void finite_kernel(const double* __restrict__ x, std::size_t count, int* __restrict__ flag) {
  const std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const std::size_t stride = static_cast<std::size_t>(gridDim.x) * blockDim.x;
  bool bad = false;
  for (std::size_t k = i; k < count; k += stride)
    if (!isfinite(x[k])) bad = true;
  if (bad) atomicOr(flag, 1);
}

__global__ void solver_failure_kernel(const rocblas_int* __restrict__ info, double* __restrict__ flag) {
  if (threadIdx.x == 0 && info[0] != 0) flag[0] += 1.0;
}

__global__ void diagonal_kernel(const double* __restrict__ A, int n, double* __restrict__ diag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) diag[i] = A[static_cast<std::size_t>(i) * n + i];
}


// ---- SP2 purification ------------------------------------------------------
// State layout (doubles): [0..1] trace of X before step k in slot k & 1,
// [2..3] freeze flag in slot 2 + (k & 1), [4] lo, [5] hi, [6] applied
// steps, [7] degenerate interval, [8 + k] idempotency error of step k.
// Every block of a step reads the slots of k and block 0 writes those of
// k + 1, so no block reads what another block of the same launch writes.
constexpr int kSp2Cols = 64, kSp2Lanes = 4, kSp2Rows = 256;

// partial[y * n + c] = sum over rows r of chunk y, r != c, of |F[r][c]|.
__global__ __launch_bounds__(kSp2Cols* kSp2Lanes) void sp2_radius_kernel(const double* __restrict__ F, int n,
                                                                         double* __restrict__ partial) {
  __shared__ double sh[kSp2Lanes][kSp2Cols];
  const int cl = threadIdx.x % kSp2Cols, lane = threadIdx.x / kSp2Cols;
  const int c = blockIdx.x * kSp2Cols + cl;
  const int r0 = blockIdx.y * kSp2Rows, r1 = min(n, r0 + kSp2Rows);
  double a = 0.0;
  if (c < n)
    for (int r = r0 + lane; r < r1; r += kSp2Lanes)
      if (r != c) a += fabs(F[static_cast<std::size_t>(r) * n + c]);
  sh[lane][cl] = a;
  __syncthreads();
  if (lane == 0 && c < n) {
    double t = sh[0][cl];
    for (int l = 1; l < kSp2Lanes; ++l) t += sh[l][cl];
    partial[static_cast<std::size_t>(blockIdx.y) * n + c] = t;
  }
}

// Gershgorin interval [lo, hi] of F and tr X0 into the state.
__global__ __launch_bounds__(kReduceThreads) void sp2_bounds_kernel(const double* __restrict__ F, int n,
                                                                    int chunks, const double* __restrict__ partial,
                                                                    double* __restrict__ state) {
  __shared__ double sh[kReduceThreads];
  const int tid = threadIdx.x;
  double lo = INFINITY, hi = -INFINITY, tr = 0.0;
  for (int c = tid; c < n; c += kReduceThreads) {
    double rad = 0.0;
    for (int y = 0; y < chunks; ++y) rad += partial[static_cast<std::size_t>(y) * n + c];
    const double d = F[static_cast<std::size_t>(c) * n + c];
    lo = fmin(lo, d - rad);
    hi = fmax(hi, d + rad);
    tr += d;
  }
  sh[tid] = lo;
  __syncthreads();
  for (int off = kReduceThreads / 2; off > 0; off >>= 1) {
    if (tid < off) sh[tid] = fmin(sh[tid], sh[tid + off]);
    __syncthreads();
  }
  lo = sh[0];
  __syncthreads();
  sh[tid] = hi;
  __syncthreads();
  for (int off = kReduceThreads / 2; off > 0; off >>= 1) {
    if (tid < off) sh[tid] = fmax(sh[tid], sh[tid + off]);
    __syncthreads();
  }
  hi = sh[0];
  __syncthreads();
  tr = block_sum(tr, sh);
  if (tid == 0) {
    const double width = hi - lo;
    const bool degenerate = !(width > 1e-12 * fmax(1.0, fmax(fabs(lo), fabs(hi))));
    state[0] = degenerate ? 0.0 : (n * hi - tr) / width;
    state[1] = 0.0;
    state[2] = degenerate ? 1.0 : 0.0;
    state[3] = 0.0;
    state[4] = lo;
    state[5] = hi;
    state[6] = 0.0;
    state[7] = degenerate ? 1.0 : 0.0;
  }
}

__global__ void sp2_init_kernel(const double* __restrict__ F, int n, const double* __restrict__ state,
                                double* __restrict__ X) {
  const std::size_t e = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= static_cast<std::size_t>(n) * n) return;
  const double lo = state[4], hi = state[5];
  const double width = state[7] != 0.0 ? 1.0 : hi - lo;
  const bool diag = e / n == e % n;
  X[e] = ((diag ? hi : 0.0) - F[e]) / width;
}

constexpr int kSp2StepThreads = 256;

__global__ __launch_bounds__(kSp2StepThreads) void sp2_step_kernel(double* __restrict__ X,
                                                                   const double* __restrict__ P, int n,
                                                                   double nocc, double* __restrict__ state,
                                                                   int k) {
  __shared__ double sh[kSp2StepThreads];
  const int tid = threadIdx.x;
  const int cur = k & 1, nxt = cur ^ 1;
  if (state[2 + cur] != 0.0) {
    if (blockIdx.x == 0 && tid == 0) {
      state[nxt] = state[cur];
      state[2 + nxt] = 1.0;
      state[8 + k] = 0.0;
    }
    return;
  }
  // tr P in a fixed order (identical in every block).
  double t = 0.0;
  for (int i = tid; i < n; i += kSp2StepThreads) t += P[static_cast<std::size_t>(i) * n + i];
  sh[tid] = t;
  __syncthreads();
  for (int off = kSp2StepThreads / 2; off > 0; off >>= 1) {
    if (tid < off) sh[tid] += sh[tid + off];
    __syncthreads();
  }
  const double trP = sh[0];
  const double trX = state[cur];
  const double err = trX - trP;
  const bool square = fabs(trP - nocc) < fabs(2.0 * trX - trP - nocc);
  // Rounding floor: the error has stopped falling over a double step while
  // already small (in the quadratic regime it falls far below half).
  const bool freeze = k >= 2 && (err <= 0.0 || (err < 1e-6 && err >= 0.5 * state[8 + k - 2]));
  const std::size_t count = static_cast<std::size_t>(n) * n;
  for (std::size_t e = static_cast<std::size_t>(blockIdx.x) * kSp2StepThreads + tid; e < count;
       e += static_cast<std::size_t>(gridDim.x) * kSp2StepThreads)
    X[e] = square ? P[e] : 2.0 * X[e] - P[e];
  if (blockIdx.x == 0 && tid == 0) {
    state[nxt] = square ? trP : 2.0 * trX - trP;
    state[2 + nxt] = freeze ? 1.0 : 0.0;
    state[6] = static_cast<double>(k + 1);
    state[8 + k] = err;
  }
}

constexpr int kCholThreads = 1024;
constexpr int kCholMaxLds = 8000;  // doubles of dynamic LDS (d[n], pivot row gp[r]), under 64 KB

__global__ __launch_bounds__(kCholThreads) void pivoted_cholesky_kernel(const double* __restrict__ P, int n, int r,
                                                                        double* __restrict__ G,
                                                                        int* __restrict__ status,
                                                                        const int* __restrict__ done) {
  // A factor already made from the previous one (done[0] == 0) stands.
  if (done && done[0] == 0) {
    if (threadIdx.x == 0) status[0] = 0;
    return;
  }
  extern __shared__ double dyn[];
  double* d = dyn;       // remaining diagonal
  double* gp = dyn + n;  // row `pivot` of G, columns < k
  __shared__ double wave_v[kCholThreads / 64];
  __shared__ int wave_i[kCholThreads / 64];
  __shared__ int pivot;
  __shared__ double pivot_v;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int i = tid; i < n; i += kCholThreads) d[i] = P[static_cast<std::size_t>(i) * n + i];
  __syncthreads();
  for (int k = 0; k < r; ++k) {
    // Arg-max of d (lowest index on ties).
    double bv = -INFINITY;
    int bi = n;
    for (int i = tid; i < n; i += kCholThreads)
      if (d[i] > bv) {
        bv = d[i];
        bi = i;
      }
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_xor(bv, off);
      const int oi = __shfl_xor(bi, off);
      if (ov > bv || (ov == bv && oi < bi)) {
        bv = ov;
        bi = oi;
      }
    }
    if (lane == 0) {
      wave_v[wave] = bv;
      wave_i[wave] = bi;
    }
    __syncthreads();
    if (tid == 0) {
      double v = wave_v[0];
      int i0 = wave_i[0];
      for (int w = 1; w < kCholThreads / 64; ++w)
        if (wave_v[w] > v || (wave_v[w] == v && wave_i[w] < i0)) {
          v = wave_v[w];
          i0 = wave_i[w];
        }
      pivot = i0;
      pivot_v = v;
    }
    __syncthreads();
    const int p = pivot;
    const double pv = pivot_v;
    if (!(pv >= 1e-8) || p >= n) {
      if (tid == 0) status[0] = 1;
      return;
    }
    for (int j = tid; j < k; j += kCholThreads) gp[j] = G[p + static_cast<std::size_t>(j) * n];
    __syncthreads();
    const double scale = 1.0 / sqrt(pv);
    for (int i = tid; i < n; i += kCholThreads) {
      double v = P[i + static_cast<std::size_t>(p) * n];
      for (int j = 0; j < k; ++j) v -= G[i + static_cast<std::size_t>(j) * n] * gp[j];
      v *= scale;
      G[i + static_cast<std::size_t>(k) * n] = v;
      d[i] = i == p ? 0.0 : d[i] - v * v;
    }
    __syncthreads();
  }
  if (tid == 0) status[0] = 0;
}

// Factor check: status[0] = 0 when the Cholesky of A^T A succeeded (info 0)
// and its diagonal spans at most `spread` (for a projected previous factor,
// kProjectedSpread: it lost no direction of the new projector), else 1.
constexpr double kProjectedSpread = 1e3;
__global__ void projected_factor_check_kernel(const rocblas_int* __restrict__ info, const double* __restrict__ R, int r,
                                              int* __restrict__ status, double spread) {
  __shared__ double lo[256], hi[256];
  double a = INFINITY, b = 0.0;
  for (int i = threadIdx.x; i < r; i += blockDim.x) {
    const double d = fabs(R[static_cast<std::size_t>(i) * r + i]);
    a = fmin(a, d);
    b = fmax(b, d);
  }
  lo[threadIdx.x] = a;
  hi[threadIdx.x] = b;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      lo[threadIdx.x] = fmin(lo[threadIdx.x], lo[threadIdx.x + off]);
      hi[threadIdx.x] = fmax(hi[threadIdx.x], hi[threadIdx.x + off]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    status[0] = info[0] == 0 && lo[0] > 0.0 && isfinite(hi[0]) && hi[0] <= spread * lo[0] ? 0 : 1;
}

// out[0] = 0 when both checks passed, else 1.
__global__ void combine_status_kernel(const int* __restrict__ checks, int* __restrict__ out) {
  if (threadIdx.x == 0) out[0] = checks[0] == 0 && checks[1] == 0 ? 0 : 1;
}

// Fixed pseudo-random sketch entries in [-1, 1): SplitMix64 of the element
// index, so every solve of a shape uses the same sketch.
__global__ void sketch_kernel(double* __restrict__ omega, std::size_t count) {
  const std::size_t stride = static_cast<std::size_t>(gridDim.x) * blockDim.x;
  for (std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride) {
    uint64_t z = static_cast<uint64_t>(i) + 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    omega[i] = static_cast<double>(z >> 11) * 0x1.0p-52 - 1.0;
  }
}

__global__ void copy_if_accepted_kernel(double* __restrict__ dst, const double* __restrict__ src, std::size_t count,
                                        const int* __restrict__ status) {
  if (status[0] != 0) return;
  const std::size_t stride = static_cast<std::size_t>(gridDim.x) * blockDim.x;
  for (std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride)
    dst[i] = src[i];
}

__global__ void copy_kernel(double* __restrict__ dst, const double* __restrict__ src, std::size_t count) {
  const std::size_t stride = static_cast<std::size_t>(gridDim.x) * blockDim.x;
  for (std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride)
    dst[i] = src[i];
}

unsigned blocks_for(std::size_t count, unsigned threads = 256) {
  return static_cast<unsigned>((count + threads - 1) / threads);
}

}  // namespace

void copy_device(double* dst, const double* src, std::size_t count, hipStream_t s) {
  if (!count) return;
  hipLaunchKernelGGL(copy_kernel, dim3(std::min<unsigned>(blocks_for(count), 8192u)), dim3(256), 0, s, dst, src, count);
  VQC_LAUNCH("copy_kernel");
}

void assemble_fock(const double* H, const double* J, const double* K, double cj, double ck, std::size_t count,
                   double* F, hipStream_t s) {
  hipLaunchKernelGGL(assemble_fock_kernel, dim3(blocks_for(count)), dim3(256), 0, s, H, J, K, cj, ck, count, F);
  VQC_LAUNCH("assemble_fock_kernel");
}

void dot_pair(const double* D, const double* A, const double* B, std::size_t count, double* out, hipStream_t s) {
  hipLaunchKernelGGL(dot_pair_kernel, dim3(1), dim3(kReduceThreads), 0, s, D, A, B, count, out);
  VQC_LAUNCH("dot_pair_kernel");
}

void rms_difference(const double* a, const double* b, std::size_t count, double* out, hipStream_t s) {
  hipLaunchKernelGGL(rms_kernel, dim3(1), dim3(kReduceThreads), 0, s, a, b, count, out);
  VQC_LAUNCH("rms_kernel");
}

void antisymmetrise(const double* T, int n, double* R, hipStream_t s) {
  const std::size_t e = static_cast<std::size_t>(n) * n;
  hipLaunchKernelGGL(antisymmetrise_kernel, dim3(blocks_for(e)), dim3(256), 0, s, T, n, R);
  VQC_LAUNCH("antisymmetrise_kernel");
}

void diis_gram_row(const double* errors, std::size_t count, int row, int m, const int* slot_of, int cap,
                   double* gram, hipStream_t s) {
  hipLaunchKernelGGL(diis_gram_kernel, dim3(m), dim3(kReduceThreads), 0, s, errors, count, row, slot_of, m, cap,
                     gram);
  VQC_LAUNCH("diis_gram_kernel");
}

void diis_solve(const double* gram, int m, const int* slot_of, int cap, double* coef, hipStream_t s) {
  hipLaunchKernelGGL(diis_solve_kernel, dim3(1), dim3(64), 0, s, gram, m, slot_of, cap, coef);
  VQC_LAUNCH("diis_solve_kernel");
}

void diis_extrapolate(const double* focks, std::size_t count, int m, const int* slot_of, const double* coef,
                      double* F, hipStream_t s) {
  hipLaunchKernelGGL(diis_extrapolate_kernel, dim3(blocks_for(count)), dim3(256), 0, s, focks, count, m, slot_of,
                     coef, F);
  VQC_LAUNCH("diis_extrapolate_kernel");
}

void scale_columns_inv_sqrt(const double* V, const double* w, int n, double* W, hipStream_t s) {
  const std::size_t e = static_cast<std::size_t>(n) * n;
  hipLaunchKernelGGL(scale_columns_kernel, dim3(blocks_for(e)), dim3(256), 0, s, V, w, n, W);
  VQC_LAUNCH("scale_columns_kernel");
}

void occupied_from_orbitals(const double* C, int n, int nocc, int ldc, double scale, double* Cocc, hipStream_t s) {
  const std::size_t e = static_cast<std::size_t>(n) * ldc;
  hipLaunchKernelGGL(occupied_kernel, dim3(blocks_for(e)), dim3(256), 0, s, C, n, nocc, ldc, scale, Cocc);
  VQC_LAUNCH("occupied_kernel");
}

void factor_from_eigen(const double* V, const double* w, const int* idx, int n, int r, int ldc, double* Cocc,
                       hipStream_t s) {
  const std::size_t e = static_cast<std::size_t>(n) * ldc;
  hipLaunchKernelGGL(factor_kernel, dim3(blocks_for(e)), dim3(256), 0, s, V, w, idx, n, r, ldc, Cocc);
  VQC_LAUNCH("factor_kernel");
}

void check_finite(const double* x, std::size_t count, int* flag, hipStream_t s) {
  hipLaunchKernelGGL(finite_kernel, dim3(std::min<unsigned>(blocks_for(count), 4096u)), dim3(256), 0, s, x, count,
                     flag);
  VQC_LAUNCH("finite_kernel");
}

std::size_t sp2_partial_size(int n) {
  return static_cast<std::size_t>((n + kSp2Rows - 1) / kSp2Rows) * n;
}

void sp2_begin(const double* F, int n, double* partial, double* state, double* X, hipStream_t s) {
  const int chunks = (n + kSp2Rows - 1) / kSp2Rows;
  hipLaunchKernelGGL(sp2_radius_kernel, dim3((n + kSp2Cols - 1) / kSp2Cols, chunks), dim3(kSp2Cols * kSp2Lanes), 0,
                     s, F, n, partial);
  VQC_LAUNCH("sp2_radius_kernel");
  hipLaunchKernelGGL(sp2_bounds_kernel, dim3(1), dim3(kReduceThreads), 0, s, F, n, chunks, partial, state);
  VQC_LAUNCH("sp2_bounds_kernel");
  const std::size_t e = static_cast<std::size_t>(n) * n;
  hipLaunchKernelGGL(sp2_init_kernel, dim3(blocks_for(e)), dim3(256), 0, s, F, n, state, X);
  VQC_LAUNCH("sp2_init_kernel");
}

void sp2_step(double* X, const double* P, int n, double nocc, double* state, int k, hipStream_t s) {
  if (k < 0 || k >= kSp2MaxSteps) throw std::invalid_argument("SP2 step index out of range");
  const std::size_t e = static_cast<std::size_t>(n) * n;
  const unsigned blocks = std::min<unsigned>(blocks_for(e, kSp2StepThreads * 4), 1024u);
  hipLaunchKernelGGL(sp2_step_kernel, dim3(std::max(1u, blocks)), dim3(kSp2StepThreads), 0, s, X, P, n, nocc, state,
                     k);
  VQC_LAUNCH("sp2_step_kernel");
}

Sp2Summary sp2_summary(const double* h, int steps) {
  Sp2Summary out;
  out.trace = h[steps & 1];
  out.converged = h[2 + (steps & 1)] != 0.0 && h[7] == 0.0;
  out.applied = static_cast<int>(h[6]);
  out.degenerate = h[7] != 0.0;
  return out;
}

bool pivoted_cholesky_fits(int n, int r) { return r >= 1 && r <= n && n + r <= kCholMaxLds; }

void pivoted_cholesky(const double* P, int n, int r, double* G, int* status, hipStream_t s, const int* done) {
  if (!pivoted_cholesky_fits(n, r)) throw std::invalid_argument("pivoted Cholesky size out of range");
  const std::size_t lds = static_cast<std::size_t>(n + r) * sizeof(double);
  hipLaunchKernelGGL(pivoted_cholesky_kernel, dim3(1), dim3(kCholThreads), lds, s, P, n, r, G, status, done);
  VQC_LAUNCH("pivoted_cholesky_kernel");
}

void projected_factor(rocblas_handle blas, const double* P, int n, int r, double* G, double* A, double* M,
                      rocblas_int* info, int* status, hipStream_t s) {
  const double one = 1.0, zero = 0.0;
  // A = P G, M = A^T A = R^T R, A <- A R^-1.
  VQC_BLAS(rocblas_dgemm(blas, rocblas_operation_none, rocblas_operation_none, n, r, n, &one, P, n, G, n, &zero, A, n));
  VQC_BLAS(rocblas_dgemm(blas, rocblas_operation_transpose, rocblas_operation_none, r, r, n, &one, A, n, A, n, &zero, M,
                         r));
  VQC_BLAS(rocsolver_dpotrf(blas, rocblas_fill_upper, r, M, r, info));
  VQC_BLAS(rocblas_dtrsm(blas, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                         rocblas_diagonal_non_unit, n, r, &one, M, r, A, n));
  hipLaunchKernelGGL(projected_factor_check_kernel, dim3(1), dim3(256), 0, s, info, M, r, status, kProjectedSpread);
  VQC_LAUNCH("projected_factor_check_kernel");
  const std::size_t count = static_cast<std::size_t>(n) * r;
  hipLaunchKernelGGL(copy_if_accepted_kernel, dim3(std::min<unsigned>(blocks_for(count), 4096u)), dim3(256), 0, s, G,
                     A, count, status);
  VQC_LAUNCH("copy_if_accepted_kernel");
}

void fill_sketch(double* omega, std::size_t count, hipStream_t s) {
  hipLaunchKernelGGL(sketch_kernel, dim3(std::min<unsigned>(blocks_for(count), 4096u)), dim3(256), 0, s, omega, count);
  VQC_LAUNCH("sketch_kernel");
}

void sketched_factor(rocblas_handle blas, const double* P, int n, int r, const double* omega, double* G, double* A,
                     double* M, rocblas_int* info, int* checks, int* status, hipStream_t s) {
  const double one = 1.0, zero = 0.0;
  // A = P Omega, then two Cholesky-QR passes (A^T A = R^T R, A <- A R^-1);
  // the second restores orthonormality lost to the first pass's conditioning.
  VQC_BLAS(rocblas_dgemm(blas, rocblas_operation_none, rocblas_operation_none, n, r, n, &one, P, n, omega, n, &zero, A,
                         n));
  for (int pass = 0; pass < 2; ++pass) {
    VQC_BLAS(rocblas_dgemm(blas, rocblas_operation_transpose, rocblas_operation_none, r, r, n, &one, A, n, A, n, &zero,
                           M, r));
    VQC_BLAS(rocsolver_dpotrf(blas, rocblas_fill_upper, r, M, r, info));
    VQC_BLAS(rocblas_dtrsm(blas, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                           rocblas_diagonal_non_unit, n, r, &one, M, r, A, n));
    // First pass: A's condition (R's diagonal spread) at most 1e6; second:
    // A is then orthonormal to rounding, so R is the identity to rounding.
    hipLaunchKernelGGL(projected_factor_check_kernel, dim3(1), dim3(256), 0, s, info, M, r, checks + pass,
                       pass == 0 ? 1e6 : 1.01);
    VQC_LAUNCH("projected_factor_check_kernel");
  }
  hipLaunchKernelGGL(combine_status_kernel, dim3(1), dim3(64), 0, s, checks, status);
  VQC_LAUNCH("combine_status_kernel");
  const std::size_t count = static_cast<std::size_t>(n) * r;
  hipLaunchKernelGGL(copy_if_accepted_kernel, dim3(std::min<unsigned>(blocks_for(count), 4096u)), dim3(256), 0, s, G,
                     A, count, status);
  VQC_LAUNCH("copy_if_accepted_kernel");
}

void note_solver_failure(const rocblas_int* info, double* flag, hipStream_t s) {
  hipLaunchKernelGGL(solver_failure_kernel, dim3(1), dim3(64), 0, s, info, flag);
  VQC_LAUNCH("solver_failure_kernel");
}

void extract_diagonal(const double* A, int n, double* diag, hipStream_t s) {
  hipLaunchKernelGGL(diagonal_kernel, dim3(blocks_for(n)), dim3(256), 0, s, A, n, diag);
  VQC_LAUNCH("diagonal_kernel");
}

}  // namespace vqc
