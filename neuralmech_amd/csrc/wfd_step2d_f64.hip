// double instantiation of the 1D / 2D step kernels (split by dtype so the two compile in parallel).
#include "wfd_step2d_impl.h"

namespace wfd {
template hipError_t launch_step2d<double>(const StepArgs<double>&, bool, int, bool, bool, bool, int, int, int, hipStream_t);
}  // namespace wfd
