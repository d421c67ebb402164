// float instantiation of the 1D / 2D step kernels with streaming access to u_prev, kc and u_next
// (split by dtype and access mode so the instantiations compile in parallel).
#include "wfd_step2d_impl.h"

namespace wfd {
template hipError_t launch_step2d_mode<float, true>(const StepArgs<float>&, bool, int, bool, bool, bool, int, int,
                                                    int, hipStream_t);
}  // namespace wfd
