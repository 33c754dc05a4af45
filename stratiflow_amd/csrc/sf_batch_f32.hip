// Instantiation of the batched-ensemble kernels (sf_kernels_impl.h) in FP32.
#include "sf_kernels_impl.h"

namespace sf {
SF_INST_BATCH(float)
}  // namespace sf
