// Nuclear-derivative kernels of the auxiliary metric class l=3; compiled as
// its own translation unit so the classes build in parallel.
#include "integrals_grad.hpp"

namespace vqc {
namespace detail {
VQC_DEFINE_METRIC_GRAD(3)
}  // namespace detail
}  // namespace vqc
