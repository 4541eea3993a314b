// Density-fitted Coulomb and exchange contractions on gfx950.
//
// Storage (HBM): the fitted tensor B is kept once, packed over AO pairs,
//   B[p][Q], p = mu(mu+1)/2 + nu (mu >= nu), row stride ldq = naux rounded up
//   to a multiple of 128 (zero padding), in FP64 and, for reduced precision,
//   also in FP32.  The occupied factor C is nbf x ldc (zero-padded columns).
// J (HBM-bound): rho_Q = sum_p d_p B[p][Q] then J_p = sum_Q B[p][Q] rho_Q,
//   two streaming passes over B, FP64 accumulation.
// K (MFMA-bound): U[mu][i][Q] = sum_nu B[(mu,nu)][Q] C[nu][i] (one GEMM per mu,
//   rows of B gathered from the packed layout), then K = W S W^T with
//   W = U viewed as nbf x (r ldq) and S the sign of each factor column, as a
//   split-K lower-triangular SYRK with a deterministic reduction.
#pragma once

#include "common.hpp"

namespace vqc {

enum class Precision : int { FP64 = 0, FP32 = 1, BF16X3 = 2 };

inline std::size_t round_up(std::size_t x, std::size_t m) { return (x + m - 1) / m * m; }

// Pick the i-tile width (16 * ni) for r columns; returns ni in 1..8.
int occupied_tile(int r);

// ---- J -------------------------------------------------------------------
void pack_density(const double* D, int n, double* dpacked, hipStream_t s);
// rho (ldq doubles) from packed density weights; `partial` needs
// coulomb_partial_size(...) doubles.
std::size_t coulomb_partial_size(std::size_t npair, int ldq);
void coulomb_rho(Precision prec, const void* B, const double* dpacked, std::size_t npair, int ldq, double* rho,
                 double* partial, hipStream_t s);
// rho = sum of `rows` partial rows (as written by occupied_u).
void rho_from_partials(const double* rho_part, int rows, int ldq, double* rho, hipStream_t s);
void coulomb_matrix(Precision prec, const void* B, const double* rho, std::size_t npair, int ldq, int n, double* J,
                    hipStream_t s);

// ---- K -------------------------------------------------------------------
// C64: n x ldc FP64 factor.  Prepares the precision-specific operand copies.
struct OccupiedOperands {
  int n = 0, r = 0, ldc = 0, ni = 1, ldct = 0;
  std::size_t ustride = 0;    // elements between consecutive mu rows of U (u_row_stride)
  const double* c64 = nullptr;
  float* c32 = nullptr;       // n x ldc
  void* ct_hi = nullptr;      // ldc x ldct bf16 (transposed), hi part
  void* ct_lo = nullptr;      // lo part
  // BF16X3 products per logical one: 3 (hi hi + hi lo + lo hi, about 2^-16
  // relative), or 1 (hi hi only, about 2^-8: the first iterations of a large
  // mixed SCF, while the density still changes by more than 1e-2).
  int terms = 3;
};
void prepare_occupied(Precision prec, OccupiedOperands& ops, hipStream_t s);
// Elements between the rows U[mu] and U[mu + 1] (each r x ldq).  The SYRK
// reads the same k window of every row at once; with a row stride of a large
// power of two times an odd number those reads fall on a few cache sets and
// memory channels.  ldq is a multiple of 128, so r ldq is an even number of
// 128-byte lines in FP32 and BF16; the reduced-precision layouts add one line
// (32 FP32 or 64 BF16 elements) so that consecutive rows start an odd number
// of lines apart.  FP64 keeps r ldq (the gradient reads FP64 U densely).
std::size_t u_row_stride(Precision prec, int r, int ldq);
// U[mu][i][Q] (r x ldq per mu): FP64 (fp64), FP32 (fp32), or hi/lo bfloat16
// planes (bf16x3).  With rho_part non-null the same pass writes
// rho_part[mu][Q] = sum_nu D[mu][nu] B[(mu,nu)][Q] (n x ldq, FP64).
void occupied_u(Precision prec, const void* B, const OccupiedOperands& ops, int ldq, void* U, const double* D,
                double* rho_part, hipStream_t s);
// K (n x n, FP64, both triangles) = sum_k sign_k W[mu][k] W[lam][k] where
// W[mu] = U[mu] (r ldq elements, rows ustride apart) and columns i < r_pos
// carry +1, the rest -1.  `partial` needs exchange_partial_size(n) doubles.
// `partial_capacity` (doubles) bounds the split-K slices to what `partial`
// holds (0: exchange_partial_size(n)).  `splits` > 0 sets the split-K slices
// per run (at most what the partial buffer holds) in place of the occupancy
// rule; the slices only regroup the k sum, so K changes at rounding level.
std::size_t exchange_partial_size(int n);
void exchange_syrk(Precision prec, const void* U, std::size_t ustride, int n, int r, int r_pos, int ldq, double* K,
                   double* partial, hipStream_t s, std::size_t partial_capacity = 0, int splits = 0, int terms = 3);
// The split-K slices per run that the occupancy rule picks for this precision
// and n (whole waves of resident blocks on 256 CUs), under the same cap.
int exchange_default_splits(Precision prec, int n, std::size_t partial_capacity = 0);
// K = Kref + W + W^T (n x n, FP64, both triangles written) with W[mu][lam] =
// sum_k X[mu][k] Y[lam][k] over k < r ldq, for two BF16X3 U buffers X and Y
// (hi/lo planes, rows ustride apart, as occupied_u writes them).  With X and Y
// the U of (C + Cref)/2 and C - Cref, W + W^T is K(C C^T - Cref Cref^T):
// an exchange update whose BF16X3 rounding scales with |C - Cref|.  `partial`
// as for exchange_syrk (FP32 slices, n^2 floats each).
void exchange_cross_bf16x3(const void* X, const void* Y, std::size_t ustride, int n, int r, int ldq,
                           const double* Kref, double* K, double* partial, hipStream_t s);
// The most split-K slices per run the partial buffer holds (at most 128).
int exchange_max_splits(int n, std::size_t partial_capacity = 0);
// C (row-major n x nocc) -> factor n x ldc scaled by sqrt(2), zero padded, and
// D = (that factor)(that factor)^T = 2 C C^T (n x n); both fully written.
void occupied_factor_and_density(const double* C, int n, int nocc, int ldc, double* factor, double* D,
                                 hipStream_t s);

// ---- conversions ---------------------------------------------------------
void convert_f64_to_f32(const double* in, float* out, std::size_t count, hipStream_t s);

}  // namespace vqc
