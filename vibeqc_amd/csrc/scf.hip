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

__global__ void diagonal_kernel(const double* __restrict__ A, int n, double* __restrict__ diag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) diag[i] = A[static_cast<std::size_t>(i) * n + i];
}

unsigned blocks_for(std::size_t count, unsigned threads = 256) {
  return static_cast<unsigned>((count + threads - 1) / threads);
}

}  // namespace

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

void extract_diagonal(const double* A, int n, double* diag, hipStream_t s) {
  hipLaunchKernelGGL(diagonal_kernel, dim3(blocks_for(n)), dim3(256), 0, s, A, n, diag);
  VQC_LAUNCH("diagonal_kernel");
}

}  // namespace vqc
