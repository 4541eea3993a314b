// Metric kernels for auxiliary f shells on the bra, (f|Q); compiled as its own
// translation unit so the classes build in parallel.
#include "integrals_kernels.hpp"

namespace vqc {
namespace detail {
VQC_DEFINE_METRIC(3)
}  // namespace detail
}  // namespace vqc
