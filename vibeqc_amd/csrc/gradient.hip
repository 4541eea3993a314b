// Analytic nuclear gradient of the DF-RHF energy on the device.
//
// Two-electron part (density_fitting.cpp:898-917 -> density_fitting_reverse_
// weights :574-595 -> density_fitting_metric_response, density_fitting_metric
// .cpp:50-99): with D = F S F^T (signed factor columns s_i), B = A L^-T,
//   rho_B = sum D B,  c = L^-T rho_B  (= M^-1 sum D A),
//   T_Qij = F_i^T B_Q F_j,  Y_ij = L^-T T_ij  (= M^-1 F_i^T A F_j),
// the energy 1/2 sum D (cJ J + cK K) has the response weights
//   W_A[mu nu][P] = cJ D_mu_nu c_P + cK sum_ij s_i s_j F_mu_i F_nu_j Y_Pij,
//   W_M[P][Q]    = -cJ/2 c_P c_Q - cK/2 sum_ij s_i s_j Y_Pij Y_Qij,
// and dE2/dR = sum W_A d(mu nu|P)/dR + sum W_M d(P|Q)/dR.  Every product is an
// FP64 rocBLAS GEMM / TRSM on the plan's stream; the derivative integrals are
// contracted with W_A and W_M by the kernels of integrals_grad.hpp.
//
// One-electron and nuclear parts (hf_gradient.cpp:8-18, mean_field.cpp:18-24):
//   sum D d(T + V)/dR - sum E dS/dR + dE_nuc/dR,  E = 2 sum_occ e_i C_i C_i^T.
#include <algorithm>
#include <cmath>

#include "integrals_gradient.hpp"
#include "plan.hpp"
#include "scf.hpp"

namespace vqc {
namespace {

__global__ void signed_columns_kernel(const double* __restrict__ F, int n, int ldc, int r, int r_pos,
                                      double* __restrict__ FS) {
  const std::size_t e = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= static_cast<std::size_t>(n) * ldc) return;
  const int j = static_cast<int>(e % ldc);
  FS[e] = j < r ? (j < r_pos ? F[e] : -F[e]) : 0.0;
}

// Column (j, i) of the (ldq x r^2) matrix Y scaled by s_i s_j.
__global__ void signed_pairs_kernel(const double* __restrict__ Y, int ldq, int r, int r_pos, double* __restrict__ Ys) {
  const std::size_t e = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const std::size_t total = static_cast<std::size_t>(r) * r * ldq;
  if (e >= total) return;
  const std::size_t col = e / ldq;
  const int i = static_cast<int>(col % r), j = static_cast<int>(col / r);
  const bool neg = (i < r_pos) != (j < r_pos);
  Ys[e] = neg ? -Y[e] : Y[e];
}

__global__ void rank1_kernel(const double* __restrict__ c, int na, double alpha, double* __restrict__ W) {
  const std::size_t e = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= static_cast<std::size_t>(na) * na) return;
  W[e] += alpha * c[e / na] * c[e % na];
}

// W_A[packed(mu, nu)][P] (+)= tmp[nu - nu0][mu][P] + cj D[mu][nu] c[P], mu >= nu.
__global__ void pack_weights_kernel(const double* __restrict__ tmp, int nu0, int nb, int n, int ldq,
                                    const double* __restrict__ D, const double* __restrict__ c, double cj,
                                    int accumulate, double* __restrict__ WA) {
  const std::size_t e = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const std::size_t total = static_cast<std::size_t>(nb) * n * ldq;
  if (e >= total) return;
  const int P = static_cast<int>(e % ldq);
  const std::size_t rest = e / ldq;
  const int mu = static_cast<int>(rest % n), nu = nu0 + static_cast<int>(rest / n);
  if (mu < nu) return;
  const std::size_t p = static_cast<std::size_t>(mu) * (mu + 1) / 2 + nu;
  const double w = tmp[e] + cj * D[static_cast<std::size_t>(mu) * n + nu] * c[P];
  WA[p * ldq + P] = accumulate ? WA[p * ldq + P] + w : w;
}

// Row-major F (n x ldc): out[mu][i] = F[mu][i] * e[i] for i < r, else 0.
__global__ void scale_factor_columns_kernel(const double* __restrict__ F, const double* __restrict__ e, int n, int ldc,
                                            int r, double* __restrict__ out) {
  const std::size_t k = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= static_cast<std::size_t>(n) * ldc) return;
  const int i = static_cast<int>(k % ldc);
  out[k] = i < r ? F[k] * e[i] : 0.0;
}

unsigned blocks_for(std::size_t count) { return static_cast<unsigned>((count + 255) / 256); }

}  // namespace

void DfPlan::df_two_electron_gradient(int r, int r_pos, double* grad) {
  if (!cholesky_) throw NotImplemented("the analytic DF gradient needs a full-rank (Cholesky-factored) metric");
  DeviceArray<double> WA(npair_ * ldq_), WM(static_cast<std::size_t>(naux_) * naux_);
  df_gradient_weights(r, r_pos, opts_.coulomb ? opts_.cj : 0.0, opts_.exchange ? opts_.ck : 0.0, false, WA.data(),
                      WM.data());
  contract_df_gradient(WA, WM, grad);
}

void DfPlan::contract_df_gradient(DeviceArray<double>& WA, DeviceArray<double>& WM, double* grad) {
  compute_three_center_gradient(orb_, ob_, aux_, al_, boys_.data(), WA.data(), ldq_, opts_.screening * 1e-10, grad,
                                stream_);
  WA.release();
  compute_metric_gradient(aux_, al_, boys_.data(), WM.data(), naux_, static_cast<int>(orb_.atoms.size()), grad,
                          stream_);
}

void DfPlan::df_gradient_weights(int r, int r_pos, double cj, double ck, bool accumulate, double* WA, double* WM) {
  if (!cholesky_) throw NotImplemented("the analytic DF gradient needs a full-rank (Cholesky-factored) metric");
  const int n = n_, na = naux_, lq = ldq_;
  const double one = 1.0, zero = 0.0;
  const double wm_beta = accumulate ? 1.0 : 0.0;
  const int acc = accumulate ? 1 : 0;
  // c = L^-T rho_B (padding rows stay zero).
  DeviceArray<double> c(lq);
  copy_device(c.data(), rho_.data(), lq, stream_);
  VQC_BLAS(rocblas_dtrsm(blas_, rocblas_side_left, rocblas_fill_lower, rocblas_operation_transpose,
                         rocblas_diagonal_non_unit, na, 1, &one, L_.data(), na, c.data(), lq));
  if (!accumulate) VQC_HIP(hipMemsetAsync(WM, 0, static_cast<std::size_t>(na) * na * sizeof(double), stream_));
  const std::size_t rl = static_cast<std::size_t>(r) * lq;
  if (ck != 0.0 && r > 0) {
    // T[j][i][Q] = sum_mu F[mu][j] U[mu][i][Q], then Y = L^-T T over Q.
    DeviceArray<double> Y(rl * r);
    const double* U = reinterpret_cast<const double*>(U_.data());
    VQC_BLAS(rocblas_dgemm(blas_, rocblas_operation_none, rocblas_operation_transpose, static_cast<int>(rl), r, n,
                           &one, U, static_cast<int>(rl), c64_.data(), ldc_, &zero, Y.data(), static_cast<int>(rl)));
    VQC_BLAS(rocblas_dtrsm(blas_, rocblas_side_left, rocblas_fill_lower, rocblas_operation_transpose,
                           rocblas_diagonal_non_unit, na, r * r, &one, L_.data(), na, Y.data(), lq));
    // W_M += -cK/2 Y diag(s_i s_j) Y^T.
    const double wk = -0.5 * ck;
    if (r_pos < r) {
      DeviceArray<double> Ys(rl * r);
      hipLaunchKernelGGL(signed_pairs_kernel, dim3(blocks_for(rl * r)), dim3(256), 0, stream_, Y.data(), lq, r, r_pos,
                         Ys.data());
      VQC_LAUNCH("signed_pairs_kernel");
      VQC_BLAS(rocblas_dgemm(blas_, rocblas_operation_none, rocblas_operation_transpose, na, na, r * r, &wk, Y.data(),
                             lq, Ys.data(), lq, &wm_beta, WM, na));
      sync();
    } else {
      VQC_BLAS(rocblas_dgemm(blas_, rocblas_operation_none, rocblas_operation_transpose, na, na, r * r, &wk, Y.data(),
                             lq, Y.data(), lq, &wm_beta, WM, na));
    }
    // V[nu][i][P] = sum_j s_j F[nu][j] Y[j][i][P] (into the U buffer, no longer needed).
    DeviceArray<double> FS(static_cast<std::size_t>(n) * ldc_);
    hipLaunchKernelGGL(signed_columns_kernel, dim3(blocks_for(FS.size())), dim3(256), 0, stream_, c64_.data(), n, ldc_,
                       r, r_pos, FS.data());
    VQC_LAUNCH("signed_columns_kernel");
    double* V = reinterpret_cast<double*>(U_.data());
    VQC_BLAS(rocblas_dgemm(blas_, rocblas_operation_none, rocblas_operation_none, static_cast<int>(rl), n, r, &one,
                           Y.data(), static_cast<int>(rl), FS.data(), ldc_, &zero, V, static_cast<int>(rl)));
    // W_A rows in chunks of nu: tmp[nu][mu][P] = cK sum_i s_i F[mu][i] V[nu][i][P].
    const std::size_t per_nu = static_cast<std::size_t>(n) * lq;
    const int nb = static_cast<int>(std::max<std::size_t>(1, std::min<std::size_t>(n, (256ull << 20) / (per_nu * 8))));
    DeviceArray<double> tmp(per_nu * nb);
    for (int nu0 = 0; nu0 < n; nu0 += nb) {
      const int cnt = std::min(nb, n - nu0);
      VQC_BLAS(rocblas_dgemm_strided_batched(blas_, rocblas_operation_none, rocblas_operation_none, lq, n, r, &ck,
                                             V + static_cast<std::size_t>(nu0) * rl, lq, static_cast<rocblas_stride>(rl),
                                             FS.data(), ldc_, 0, &zero, tmp.data(), lq,
                                             static_cast<rocblas_stride>(per_nu), cnt));
      hipLaunchKernelGGL(pack_weights_kernel, dim3(blocks_for(per_nu * cnt)), dim3(256), 0, stream_, tmp.data(), nu0,
                         cnt, n, lq, D_.data(), c.data(), cj, acc, WA);
      VQC_LAUNCH("pack_weights_kernel");
    }
    sync();
  } else {
    // Coulomb only: W_A = cJ D c.
    const int nb = static_cast<int>(std::max<std::size_t>(1, std::min<std::size_t>(n, (256ull << 20) / (static_cast<std::size_t>(n) * lq * 8))));
    DeviceArray<double> tmp(static_cast<std::size_t>(n) * lq * nb);
    VQC_HIP(hipMemsetAsync(tmp.data(), 0, tmp.bytes(), stream_));
    for (int nu0 = 0; nu0 < n; nu0 += nb) {
      const int cnt = std::min(nb, n - nu0);
      hipLaunchKernelGGL(pack_weights_kernel, dim3(blocks_for(static_cast<std::size_t>(n) * lq * cnt)), dim3(256), 0,
                         stream_, tmp.data(), nu0, cnt, n, lq, D_.data(), c.data(), cj, acc, WA);
      VQC_LAUNCH("pack_weights_kernel");
    }
    sync();
  }
  if (cj != 0.0) {
    hipLaunchKernelGGL(rank1_kernel, dim3(blocks_for(static_cast<std::size_t>(na) * na)), dim3(256), 0, stream_,
                       c.data(), na, -0.5 * cj, WM);
    VQC_LAUNCH("rank1_kernel");
  }
  sync();
}

void DfPlan::one_electron_gradient(const double* D, const double* eps, int nocc, double* grad) {
  const int n = n_;
  const std::size_t nn = static_cast<std::size_t>(n) * n;
  const double one = 1.0, zero = 0.0;
  // E = sum_i e_i F_i F_i^T with F = sqrt(2) C_occ (c64_): 2 sum e_i C_i C_i^T.
  DeviceArray<double> Fe(static_cast<std::size_t>(n) * ldc_), E(nn);
  hipLaunchKernelGGL(scale_factor_columns_kernel, dim3(blocks_for(Fe.size())), dim3(256), 0, stream_, c64_.data(), eps, n,
                     ldc_, nocc, Fe.data());
  VQC_LAUNCH("scale_factor_columns_kernel");
  VQC_BLAS(rocblas_dgemm(blas_, rocblas_operation_transpose, rocblas_operation_none, n, n, std::max(nocc, 1), &one,
                         Fe.data(), ldc_, c64_.data(), ldc_, &zero, E.data(), n));
  compute_one_electron_gradient(orb_, ob_, boys_.data(), D, E.data(), grad, stream_);
}

std::vector<double> DfPlan::nuclear_repulsion_gradient() const {
  const std::size_t na = orb_.atoms.size();
  std::vector<double> g(na * 3, 0.0);
  for (std::size_t i = 0; i < na; ++i)
    for (std::size_t j = 0; j < na; ++j) {
      if (i == j) continue;
      const auto& a = orb_.atoms[i];
      const auto& b = orb_.atoms[j];
      const double d[3] = {a.xyz[0] - b.xyz[0], a.xyz[1] - b.xyz[1], a.xyz[2] - b.xyz[2]};
      const double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      const double f = -static_cast<double>(a.z) * b.z / (r * r * r);
      for (int x = 0; x < 3; ++x) g[i * 3 + x] += f * d[x];
    }
  return g;
}

}  // namespace vqc
