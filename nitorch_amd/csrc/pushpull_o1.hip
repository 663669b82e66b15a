// Order-1 specialisations of the 3-D float push/pull kernels (one
// translation unit per order so that the library builds in parallel).
#include "pushpull_tile.h"

namespace ni_amd {
NI_DEFINE_DISPATCH_ORDER(1)
}  // namespace ni_amd
