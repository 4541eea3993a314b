// Nuclear-derivative kernels of the orbital class (l=3, l=3); compiled as its
// own translation unit so the classes build in parallel.
#include "integrals_grad.hpp"

namespace vqc {
namespace detail {
VQC_DEFINE_GRAD_CLASS(3, 3)
}  // namespace detail
}  // namespace vqc
