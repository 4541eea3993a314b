// Device integral generation: Hcore pieces, the (P|Q) metric and the packed
// three-centre tensor (mu nu|P).
#pragma once

#include <vector>

#include "basis.hpp"
#include "common.hpp"

namespace vqc {

// Per-shell record shared by the kernels.  fn_off is the public AO offset of
// an orbital shell, or the internal auxiliary offset of an auxiliary shell.
struct DevShell {
  double x, y, z;
  int l, prim_off, nprim, fn_off;
};

// One auxiliary function: its shell, its Hermite coefficient vector and its
// internal auxiliary index.
struct AuxFunction {
  int shell, coef_off, index, pad;
};

// Boys function table F_m(T_k), T_k = k * kBoysStep, used by every kernel.
constexpr int kBoysCols = 20;
constexpr double kBoysStep = 0.1;
constexpr double kBoysTmax = 36.0;
constexpr int kBoysRows = 361;
// The table on the host, computed once per process (long-double series).
const std::vector<double>& boys_table_host();

// A basis uploaded for integral evaluation.
struct DeviceBasis {
  int representation = 0;
  int nshell = 0;
  std::size_t nfunctions = 0;
  std::vector<DevShell> host_shells;
  DeviceArray<DevShell> shells;
  DeviceArray<double> exps, coefs;
  // Per-l AO transform rows (public x raw Cartesian monomials), offset table.
  DeviceArray<double> transform;
  std::vector<int> transform_off;  // [l] -> offset
};

// The auxiliary side: functions grouped by angular momentum (internal order).
struct AuxiliaryLayout {
  DeviceBasis basis;
  std::vector<int> class_begin;      // [l], [l+1] -> range of internal functions
  DeviceArray<AuxFunction> functions;
  DeviceArray<double> hermite;       // dense Hermite coefficient vectors
  std::vector<int> public_of_internal;
};

void upload_basis(const System& sys, DeviceBasis& out, bool auxiliary_offsets, hipStream_t s);
void build_auxiliary_layout(const System& aux, AuxiliaryLayout& out, hipStream_t s);

// S, T, V (nbf x nbf, row-major, public AO order).
void compute_one_electron(const System& sys, const DeviceBasis& basis, const double* boys, double* S,
                          double* T, double* V, hipStream_t s);

// (P|Q) in internal auxiliary order, naux x naux row-major.
void compute_metric(const AuxiliaryLayout& aux, const double* boys, double* M, hipStream_t s);

// Packed (mu nu|P): row p = mu(mu+1)/2 + nu (mu >= nu), row stride ld >= naux
// (internal auxiliary order).  The npair x ld buffer is fully overwritten
// (screened entries and padding columns are zero).  With a shell range, only
// the rows of the pairs whose larger shell index lies in [shell_begin,
// shell_end) are computed: rows f0(f0+1)/2 up to f1(f1+1)/2, f0 and f1 the
// first AO functions of shell_begin and shell_end, stored from A's first row.
void compute_three_center(const System& orb, const DeviceBasis& basis, const AuxiliaryLayout& aux,
                          const double* boys, double* A, std::size_t ld, double screening, hipStream_t s,
                          int shell_begin = 0, int shell_end = -1);

}  // namespace vqc
