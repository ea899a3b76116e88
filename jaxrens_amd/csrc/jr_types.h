// Plain data shared by the library's translation units and its kernels:
// the device move table entry, the population pointers and the
// device-resident iteration state.
#pragma once

namespace jr {

// work-queue counters: [0] single-launch walks, [1 + j] walk of size class j
constexpr int JR_NQUEUE = 4;

struct DevMove {
  int type, n_reflect, n_leapfrog, flat_v_prior, n_species, species_mask;
  double perturb_angle, max_vpa, min_vpa, min_aspect;
  int refresh_direction, dir_slot;  // gmc
};

// Device-resident iteration state, read by every kernel of an iteration so a
// captured graph can be replayed without parameter updates.
struct IterState {
  long long i;         // outer-loop iteration index (RE firing rule)
  long long rng_iter;  // monotone counter feeding the Philox counter
  int block_pos;       // slot in the per-block output buffers
  int re_interval;
};

struct PopPtrs {
  double* pos;
  int* types;
  int* n_atoms;
  double* cell;
  double* energy;
  double* dir;
  // directions of the species-scoped GMC moves, n_xslots slots of xstride
  // doubles ([R*K][Nmax][3] each); they travel with the walker
  double* dirx = nullptr;
  int n_xslots = 0;
  size_t xstride = 0;
};

}  // namespace jr
