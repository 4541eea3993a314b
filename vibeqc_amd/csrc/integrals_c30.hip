// Integral kernels of the orbital class (l=3, l=0) and the metric kernels for
// auxiliary f shells on the bra, (f|Q); compiled as its own translation unit
// so the classes build in parallel.
#include "integrals_kernels.hpp"

namespace vqc {
namespace detail {
VQC_DEFINE_CLASS(3, 0)
VQC_DEFINE_METRIC(3)
}  // namespace detail
}  // namespace vqc
