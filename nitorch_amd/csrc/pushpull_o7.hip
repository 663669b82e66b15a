// Order-7 specialisations of the 3-D float push/pull kernels (one
// translation unit per order so that the library builds in parallel).
#include "pushpull_tile.h"

namespace ni_amd {
NI_DEFINE_DISPATCH_ORDER(7)
}  // namespace ni_amd
