// Small device kernels of the RHF SCF loop (all FP64).
#pragma once

#include "common.hpp"

namespace vqc {

// dst[i] = src[i] for i < count by a kernel on stream s.  Device-to-device
// copies between the plan's buffers go through this rather than
// hipMemcpyAsync, so every producer and consumer of a buffer is an ordinary
// kernel of the same stream.
void copy_device(double* dst, const double* src, std::size_t count, hipStream_t s);

// F = H + cj J + ck K.
void assemble_fock(const double* H, const double* J, const double* K, double cj, double ck, std::size_t count,
                   double* F, hipStream_t s);
// out[0] = sum D .* A, out[1] = sum D .* B (both partial-sum reduced on device).
void dot_pair(const double* D, const double* A, const double* B, std::size_t count, double* out, hipStream_t s);
// out = sqrt(mean((a - b)^2)).
void rms_difference(const double* a, const double* b, std::size_t count, double* out, hipStream_t s);
// R = T - T^T (n x n).
void antisymmetrise(const double* T, int n, double* R, hipStream_t s);
// gram[row * cap + j] = gram[j * cap + row] = <E_row, E_j> for j < m (history slots).
void diis_gram_row(const double* errors, std::size_t count, int row, int m, const int* slot_of, int cap,
                   double* gram, hipStream_t s);
// Solve the augmented Pulay system for m entries (order given by slot_of);
// coef[j] for j < m; a singular system returns coef = e_{m-1}.
void diis_solve(const double* gram, int m, const int* slot_of, int cap, double* coef, hipStream_t s);
// F = sum_j coef[j] * focks[slot_of[j]].
void diis_extrapolate(const double* focks, std::size_t count, int m, const int* slot_of, const double* coef,
                      double* F, hipStream_t s);
// Column scale: W[:, j] = V[:, j] * f(w[j]) with f = w^-1/2 (column-major n x n).
void scale_columns_inv_sqrt(const double* V, const double* w, int n, double* W, hipStream_t s);
// Cocc (row-major n x ldc) = scale * C[:, :nocc] (C column-major n x n), zero padded.
void occupied_from_orbitals(const double* C, int n, int nocc, int ldc, double scale, double* Cocc, hipStream_t s);
// Cocc (row-major n x ldc) from selected eigenvectors: column i <- sqrt|w[idx[i]]| V[:, idx[i]].
void factor_from_eigen(const double* V, const double* w, const int* idx, int n, int r, int ldc, double* Cocc,
                       hipStream_t s);
// Non-finite detector: flag[0] |= any non-finite in x.
void check_finite(const double* x, std::size_t count, int* flag, hipStream_t s);
// Eigensolver status: flag[0] += 1 when info[0] != 0 (a rocSOLVER routine
// did not converge), so a later download of flag reports the failure with no
// extra synchronisation.
void note_solver_failure(const rocblas_int* info, double* flag, hipStream_t s);
// Symmetric copy from a column-major Cholesky factor: diag[i] = L[i + i*n].
void extract_diagonal(const double* A, int n, double* diag, hipStream_t s);

// Second-order spectral projection (SP2 purification; A. M. N. Niklasson,
// Phys. Rev. B 66, 155115 (2002)) of a symmetric F (n x n, orthonormal
// basis): X0 = (hi I - F) / (hi - lo) over F's Gershgorin interval, then
// X <- X^2 or 2X - X^2, whichever brings tr X closer to nocc, until the
// idempotency error tr(X - X^2) reaches its rounding floor.  X converges to
// the projector onto the nocc lowest eigenvectors of F.  The state array
// (kSp2StateSize doubles) carries the traces, the freeze flag and the
// error history on the device, so steps need no host round trip.
constexpr int kSp2MaxSteps = 100;
constexpr int kSp2StateSize = 8 + kSp2MaxSteps;
std::size_t sp2_partial_size(int n);
void sp2_begin(const double* F, int n, double* partial, double* state, double* X, hipStream_t s);
// Step k (0-based) given P = X X; the step after the error floor is reached
// and every later one leave X unchanged.
void sp2_step(double* X, const double* P, int n, double nocc, double* state, int k, hipStream_t s);
// state[0..7] after `steps` steps: trace, converged flag, applied steps.
struct Sp2Summary {
  double trace;
  bool converged;
  int applied;
  bool degenerate;
};
Sp2Summary sp2_summary(const double* host_state, int steps);
// G (n x r, column-major, leading dimension n) with P = G G^T for a
// symmetric positive semidefinite P of rank r, such as a converged SP2
// projector: r steps of Cholesky with diagonal pivoting in one workgroup
// (pivot: largest remaining diagonal, lowest index on ties).  status[0] = 1
// when a pivot falls below 1e-8, i.e. P is not of rank r.  Requires
// pivoted_cholesky_fits(n, r).
bool pivoted_cholesky_fits(int n, int r);
// With `done` non-null and done[0] == 0 (a factor was already accepted by
// projected_factor), status[0] = 0 and G is left as it is.
void pivoted_cholesky(const double* P, int n, int r, double* G, int* status, hipStream_t s, const int* done = nullptr);
// Factor of a projector P (rank r) from the previous iteration's factor G
// (n x r, column-major), when its projection still spans P's range:
// A = P G, A^T A = R^T R, G <- A R^-1 (so G G^T = P).  status[0] = 0 when
// accepted (Cholesky info 0 and R's diagonal within a factor 1e3), else 1 and
// G is unchanged.  A (n x r) and M (r x r) are workspaces; everything stays on
// the stream.
void projected_factor(rocblas_handle blas, const double* P, int n, int r, double* G, double* A, double* M,
                      rocblas_int* info, int* status, hipStream_t s);
// Fixed pseudo-random entries in [-1, 1) (the same for every call).
void fill_sketch(double* omega, std::size_t count, hipStream_t s);
// Factor of a projector P (rank r) from a random sketch of its range:
// A = P Omega (Omega n x r from fill_sketch), then two Cholesky-QR passes, so
// G G^T = P (to rounding) without the previous factor.  status[0] = 0 when
// accepted (both Cholesky factorisations succeeded, A's condition at most
// 1e6 and orthonormal after the second pass), else 1 and G is unchanged.
// A, M as projected_factor; `checks` holds two ints.
void sketched_factor(rocblas_handle blas, const double* P, int n, int r, const double* omega, double* G, double* A,
                     double* M, rocblas_int* info, int* checks, int* status, hipStream_t s);

}  // namespace vqc
