// C-ABI shim over the reference's own basis normalisation, AO expansions and
// metric factorisation (src/molecule/basis.cpp, src/integrals/
// density_fitting_metric.cpp:101-152), built from the reference sources where
// they lie by oracle/Makefile into oracle/_ref/libref.so.  Test
// infrastructure only: the tests use it to check the oracle's restatements
// (oracle.normalise_shell, oracle.spherical_rows, oracle.cart_components,
// oracle.metric_inverse_sqrt) against the reference code itself.
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "core/types.hpp"
#include "integrals/density_fitting_metric.hpp"
#include "molecule/basis.hpp"

extern "C" int ref_factor_metric(const double* metric, int n, double relative_threshold, double* inverse_sqrt,
                                 int* rank, double* condition) {
  try {
    const std::vector<double> m(metric, metric + static_cast<std::size_t>(n) * n);
    const auto f = generativeqc::integrals::factor_density_fitting_metric(m, static_cast<std::size_t>(n),
                                                                          relative_threshold);
    std::memcpy(inverse_sqrt, f.inverse_square_root.data(), f.inverse_square_root.size() * sizeof(double));
    *rank = static_cast<int>(f.effective_rank);
    *condition = f.condition_number;
    return 0;
  } catch (const std::exception&) {
    return 1;
  }
}

// Public AOs of a shell as rows over its Cartesian components (CCA order):
// out[row * cols + component].  Returns the row count, or -1.
extern "C" int ref_ao_expansion(unsigned l, int spherical, double* out, int max_rows, int cols) {
  try {
    const auto rep = spherical ? GENERATIVEQC_BASIS_SPHERICAL : GENERATIVEQC_BASIS_CARTESIAN;
    const auto rows = generativeqc::molecule::ao_expansions(l, rep);
    const auto comps = generativeqc::molecule::cartesian_components(l);
    if (static_cast<int>(rows.size()) > max_rows || static_cast<int>(comps.size()) != cols) return -1;
    std::memset(out, 0, sizeof(double) * static_cast<std::size_t>(max_rows) * cols);
    for (std::size_t r = 0; r < rows.size(); ++r)
      for (const auto& term : rows[r])
        for (int c = 0; c < cols; ++c)
          if (comps[c] == term.component) out[r * cols + c] += term.coefficient;
    return static_cast<int>(rows.size());
  } catch (const std::exception&) {
    return -1;
  }
}

// Cartesian components of angular momentum l in the reference's order:
// out[3 c + axis].  Returns the count, or -1.
extern "C" int ref_cartesian_components(unsigned l, int* out, int max_count) {
  const auto comps = generativeqc::molecule::cartesian_components(l);
  if (static_cast<int>(comps.size()) > max_count) return -1;
  for (std::size_t c = 0; c < comps.size(); ++c)
    for (int a = 0; a < 3; ++a) out[3 * c + a] = static_cast<int>(comps[c][a]);
  return static_cast<int>(comps.size());
}

// Normalised primitive coefficients of one contracted shell on a helium atom
// (the validator's radial factor and contraction normalisation).  0 on success.
extern "C" int ref_normalize_shell(unsigned l, int nprim, const double* exps, const double* coefs, double* out) {
  generativeqc::core::System sys;
  sys.atoms.push_back({2, {0.0, 0.0, 0.0}, 0});
  generativeqc::core::Shell sh;
  sh.atom_index = 0;
  sh.angular_momentum = l;
  for (int p = 0; p < nprim; ++p) sh.primitives.push_back({exps[p], coefs[p]});
  sys.shells.push_back(sh);
  sys.charge = 0;
  sys.multiplicity = 1;
  sys.basis_representation = GENERATIVEQC_BASIS_CARTESIAN;
  std::string detail;
  if (generativeqc::molecule::validate_and_normalize(sys, detail) != GENERATIVEQC_STATUS_SUCCESS) return 1;
  for (int p = 0; p < nprim; ++p) out[p] = sys.shells[0].primitives[p].coefficient;
  return 0;
}
