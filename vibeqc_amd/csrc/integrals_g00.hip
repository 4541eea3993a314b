// Nuclear-derivative kernels of the orbital class (l=0, l=0); compiled as its
// own translation unit so the classes build in parallel.
#include "integrals_grad.hpp"

namespace vqc {
namespace detail {
VQC_DEFINE_GRAD_CLASS(0, 0)
VQC_DEFINE_METRIC_GRAD(0)
}  // namespace detail
}  // namespace vqc
