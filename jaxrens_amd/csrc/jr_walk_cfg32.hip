// Walk-kernel instantiations for kernel configuration 32 (bit meanings in
// jr_launch.h), in a translation unit of its own so the library builds in
// parallel.
#include "jr_walk_inst.h"

JR_WALK_INSTANTIATE(32)
