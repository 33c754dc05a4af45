// Instantiation of the single-run kernels (sf_kernels_impl.h) in FP32, with the
// host-side helpers declared in sf_kernels.h.
#define SF_KERNELS_HOST_DEFS
#include "sf_kernels_impl.h"

namespace sf {
SF_INST_CORE(float)
}  // namespace sf
