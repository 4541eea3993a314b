// Integral kernels of the orbital class (l=3, l=3); compiled as its own
// translation unit so the classes build in parallel.
#include "integrals_kernels.hpp"

namespace vqc {
namespace detail {
VQC_DEFINE_CLASS(3, 3)
}  // namespace detail
}  // namespace vqc
