// Instantiation of the single-run kernels (sf_kernels_impl.h) in FP64.
#include "sf_kernels_impl.h"

namespace sf {
SF_INST_CORE(double)
}  // namespace sf
