// Definitions of the walk-kernel launchers declared in jr_launch.h; included
// only by the translation units that instantiate them (jr_walk_cfg*.hip).
#pragma once

#include "jr_launch.h"

namespace jr {

template <int A, int CFG, int OCC>
void launch_walk_t(const WalkArgs& a, int blocks, int wpb, size_t lds, hipStream_t s) {
  hipLaunchKernelGGL((walk_kernel<A, CFG, OCC>), dim3(blocks), dim3(64 * wpb), lds, s, a);
}

template <int CFG>
void launch_walk_cfg(int A, int occ, const WalkArgs& a, int blocks, int wpb, size_t lds,
                     hipStream_t s) {
  if (A == 1) {
    if constexpr (!(CFG & 60)) {
      if (occ == 3) {
        launch_walk_t<1, CFG & 3, 3>(a, blocks, wpb, lds, s);
        return;
      }
    }
    launch_walk_t<1, CFG, 0>(a, blocks, wpb, lds, s);
  } else if (A == 2) {
    launch_walk_t<2, CFG, 0>(a, blocks, wpb, lds, s);
  } else {
    launch_walk_t<4, CFG, 0>(a, blocks, wpb, lds, s);
  }
}

template <int CFG>
const void* walk_fn_cfg(int A, int occ) {
  if (A == 1) {
    if constexpr (!(CFG & 60)) {
      if (occ == 3) return (const void*)walk_kernel<1, CFG & 3, 3>;
    }
    return (const void*)walk_kernel<1, CFG, 0>;
  }
  if (A == 2) return (const void*)walk_kernel<2, CFG, 0>;
  return (const void*)walk_kernel<4, CFG, 0>;
}

}  // namespace jr

#define JR_WALK_INSTANTIATE(CFG)                                                        \
  template void jr::launch_walk_cfg<CFG>(int, int, const jr::WalkArgs&, int, int, size_t, \
                                         hipStream_t);                                  \
  template const void* jr::walk_fn_cfg<CFG>(int, int);
