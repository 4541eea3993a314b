// Nuclear-derivative integrals contracted with response weights; each call
// adds its contribution to grad (natom x 3, device, FP64).
#pragma once

#include "integrals.hpp"

namespace vqc {

// grad += sum_{mu nu P} W_A[mu nu][P] d(mu nu|P)/dR, with W_A packed over AO
// pairs (row p = mu(mu+1)/2 + nu, mu >= nu; symmetric in mu, nu) and internal
// auxiliary columns, row stride ld.
void compute_three_center_gradient(const System& orb, const DeviceBasis& basis, const System& aux_sys,
                                   const AuxiliaryLayout& aux, const double* boys, const double* W, std::size_t ld,
                                   double screening, double* grad, hipStream_t s);

// grad += sum_{PQ} W_M[P][Q] d(P|Q)/dR, W_M in internal auxiliary order, row
// stride ld.
void compute_metric_gradient(const System& aux_sys, const AuxiliaryLayout& aux, const double* boys, const double* W,
                             std::size_t ld, int natom, double* grad, hipStream_t s);

// grad += sum D d(T + V)/dR - sum E dS/dR (n x n symmetric public-AO D, E).
void compute_one_electron_gradient(const System& sys, const DeviceBasis& basis, const double* boys,
                                   const double* dens, const double* edens, double* grad, hipStream_t s);

}  // namespace vqc
