// Batched-evaluation kernel instantiations and the walk-kernel dispatch over
// kernel configurations (the walk kernels themselves live in
// jr_walk_cfg*.hip).
#include "jr_launch.h"

namespace jr {

namespace {

template <int A, int CFG>
void launch_eval_t(const EvalArgs& a, hipStream_t s) {
  constexpr bool team = (CFG & 32) != 0;
  const int wpb = team ? a.team : 4;
  const size_t lds = block_lds_bytes(A, wpb, a.S, team ? a.chunks : 1);
  const int blocks = team ? a.B : (a.B + wpb - 1) / wpb;
  hipLaunchKernelGGL((eval_kernel<A, CFG>), dim3(blocks), dim3(64 * wpb), lds, s, a);
}

template <int CFG>
void launch_eval_c(int A, const EvalArgs& a, hipStream_t s) {
  if (A == 1) launch_eval_t<1, CFG>(a, s);
  else if (A == 2) launch_eval_t<2, CFG>(a, s);
  else launch_eval_t<4, CFG>(a, s);
}

}  // namespace

void launch_eval(int A, int cfg, const EvalArgs& a, hipStream_t s) {
  switch (cfg) {
    case 0: launch_eval_c<0>(A, a, s); break;
    case 1: launch_eval_c<1>(A, a, s); break;
    case 2: launch_eval_c<2>(A, a, s); break;
    case 3: launch_eval_c<3>(A, a, s); break;
    case 4: launch_eval_c<4>(A, a, s); break;
    case 5: launch_eval_c<5>(A, a, s); break;
    case 32: launch_eval_c<32>(A, a, s); break;
    default: launch_eval_c<33>(A, a, s); break;
  }
}

void launch_walk(int A, int cfg, int occ, const WalkArgs& a, int blocks, int wpb,
                 size_t lds, hipStream_t s) {
  switch (cfg) {
    case 0: launch_walk_cfg<0>(A, occ, a, blocks, wpb, lds, s); break;
    case 1: launch_walk_cfg<1>(A, occ, a, blocks, wpb, lds, s); break;
    case 2: launch_walk_cfg<2>(A, occ, a, blocks, wpb, lds, s); break;
    case 3: launch_walk_cfg<3>(A, occ, a, blocks, wpb, lds, s); break;
    case 4: launch_walk_cfg<4>(A, occ, a, blocks, wpb, lds, s); break;
    case 5: launch_walk_cfg<5>(A, occ, a, blocks, wpb, lds, s); break;
    case 8: launch_walk_cfg<8>(A, occ, a, blocks, wpb, lds, s); break;
    case 9: launch_walk_cfg<9>(A, occ, a, blocks, wpb, lds, s); break;
    case 12: launch_walk_cfg<12>(A, occ, a, blocks, wpb, lds, s); break;
    case 13: launch_walk_cfg<13>(A, occ, a, blocks, wpb, lds, s); break;
    case 16: launch_walk_cfg<16>(A, occ, a, blocks, wpb, lds, s); break;
    case 17: launch_walk_cfg<17>(A, occ, a, blocks, wpb, lds, s); break;
    case 24: launch_walk_cfg<24>(A, occ, a, blocks, wpb, lds, s); break;
    case 25: launch_walk_cfg<25>(A, occ, a, blocks, wpb, lds, s); break;
    case 32: launch_walk_cfg<32>(A, occ, a, blocks, wpb, lds, s); break;
    case 33: launch_walk_cfg<33>(A, occ, a, blocks, wpb, lds, s); break;
    case 40: launch_walk_cfg<40>(A, occ, a, blocks, wpb, lds, s); break;
    default: launch_walk_cfg<41>(A, occ, a, blocks, wpb, lds, s); break;
  }
}

const void* walk_fn(int A, int cfg, int occ) {
  switch (cfg) {
    case 0: return walk_fn_cfg<0>(A, occ);
    case 1: return walk_fn_cfg<1>(A, occ);
    case 2: return walk_fn_cfg<2>(A, occ);
    case 3: return walk_fn_cfg<3>(A, occ);
    case 4: return walk_fn_cfg<4>(A, occ);
    case 5: return walk_fn_cfg<5>(A, occ);
    case 8: return walk_fn_cfg<8>(A, occ);
    case 9: return walk_fn_cfg<9>(A, occ);
    case 12: return walk_fn_cfg<12>(A, occ);
    case 13: return walk_fn_cfg<13>(A, occ);
    case 16: return walk_fn_cfg<16>(A, occ);
    case 17: return walk_fn_cfg<17>(A, occ);
    case 24: return walk_fn_cfg<24>(A, occ);
    case 25: return walk_fn_cfg<25>(A, occ);
    case 32: return walk_fn_cfg<32>(A, occ);
    case 33: return walk_fn_cfg<33>(A, occ);
    case 40: return walk_fn_cfg<40>(A, occ);
    default: return walk_fn_cfg<41>(A, occ);
  }
}

}  // namespace jr
