// Instantiation of the batched-ensemble kernels (sf_kernels_impl.h) in FP64.
#include "sf_kernels_impl.h"

namespace sf {
SF_INST_BATCH(double)
}  // namespace sf
