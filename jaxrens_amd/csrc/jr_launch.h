// Host launchers of the templated walk and evaluation kernels.  Each kernel
// configuration is instantiated in a translation unit of its own
// (jr_walk_cfg*.hip, jr_eval_inst.hip) so the library builds in parallel;
// this header only declares them.
#pragma once

#include <hip/hip_runtime.h>

#include "jr_kernels.h"

namespace jr {

// cfg bit 0: species table (S > 1); bit 1: FP32 pair terms; bit 2: image
// sums; bit 3: shear, stretch or species_swap in the move table; bit 4:
// minimum-distance constraint (bits 2-4 FP64 only, bit 4 without bit 2);
// bit 5: evaluation team (a block of eval_chunks waves per chain or
// configuration, with bits 0 and 3 only).
// A = atoms per lane (1, 2, 4); occ = 0 or 3 (the occupancy-3 build exists
// for A = 1 without bits 2-5).
template <int CFG>
void launch_walk_cfg(int A, int occ, const WalkArgs& a, int blocks, int wpb, size_t lds,
                     hipStream_t s);
template <int CFG>
const void* walk_fn_cfg(int A, int occ);

void launch_walk(int A, int cfg, int occ, const WalkArgs& a, int blocks, int wpb,
                 size_t lds, hipStream_t s);
const void* walk_fn(int A, int cfg, int occ);

// cfg bits 0-2, or 32 | bit 0
void launch_eval(int A, int cfg, const EvalArgs& a, hipStream_t s);

}  // namespace jr
