// C-ABI shim over the reference's own metric factorisation
// (src/integrals/density_fitting_metric.cpp:101-152, built from the reference
// sources where they lie by oracle/Makefile into oracle/_ref/).  Test
// infrastructure only: the tests use it to check the oracle's restatement
// (oracle.metric_inverse_sqrt) against the reference code itself.
#include <cstring>
#include <exception>
#include <vector>

#include "integrals/density_fitting_metric.hpp"

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
