// Pair-potential helpers shared by the library's translation units:
// kernel-configuration bits of a jr_pot, its validation, and the argument
// checks of the batched evaluators.
#pragma once

#include "jr_host.h"
#include "jr_launch.h"

namespace jrh {

inline int atoms_per_lane(int nmax) { return nmax <= 64 ? 1 : (nmax <= 128 ? 2 : 4); }

inline bool pot_images(const jr_pot* pot) {
  return pot->images[0] > 0 || pot->images[1] > 0 || pot->images[2] > 0;
}

inline int pot_chunks(const jr_pot* pot) { return pot->eval_chunks > 1 ? pot->eval_chunks : 1; }

// kernel configuration bits 0-2 and 5 (jr_launch.h)
inline int pot_cfg(const jr_pot* pot) {
  return (pot->n_species > 1 ? 1 : 0) |
         (pot->precision == JR_PREC_FP32_PAIRS ? 2 : 0) |
         (pot_images(pot) ? 4 : 0) | (pot_chunks(pot) > 1 ? 32 : 0);
}

// LDS one workgroup may allocate on `device` (queried once per device).
inline size_t lds_per_block(int device) {
  static size_t cache[64] = {0};
  const int d = device >= 0 && device < 64 ? device : 0;
  if (!cache[d]) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) !=
            hipSuccess ||
        v <= 0)
      v = 65536;
    cache[d] = (size_t)v;
  }
  return cache[d];
}

// Waves of an evaluation team for `chunks` chunks at A atoms per lane: the
// largest divisor of chunks up to 4 (one wave per SIMD of a CU) whose block
// LDS fits; each wave evaluates chunks / waves consecutive chunks.
inline int team_waves(int A, int chunks, int S, int device) {
  const size_t cap = lds_per_block(device);
  int w = chunks < 4 ? chunks : 4;
  while (w > 1 && jr::block_lds_bytes(A, w, S, chunks) > cap) w >>= 1;
  return w;
}

inline void set_images(int* dst, const jr_pot& pot) {
  for (int d = 0; d < 3; ++d) dst[d] = pot.images[d];
}

inline int validate_pot(const jr_pot* pot) {
  if (!pot) return set_global_err(JR_E_ARG, "pot is NULL");
  if (pot->n_species < 1 || pot->n_species > JR_MAX_SPECIES)
    return set_global_err(JR_E_ARG, "n_species must be in [1, %d], got %d",
                          JR_MAX_SPECIES, pot->n_species);
  if (!pot->eps4 || !pot->eps24 || !pot->sig2)
    return set_global_err(JR_E_ARG, "pair tables must not be NULL");
  if (pot->precision != JR_PREC_FP64 && pot->precision != JR_PREC_FP32_PAIRS)
    return set_global_err(JR_E_ARG, "unknown precision %d", pot->precision);
  for (int d = 0; d < 3; ++d)
    if (pot->images[d] < 0 || pot->images[d] > JR_MAX_IMAGES)
      return set_global_err(JR_E_ARG, "images[%d] must be in [0, %d], got %d", d,
                            JR_MAX_IMAGES, pot->images[d]);
  if (pot_images(pot) && pot->precision != JR_PREC_FP64)
    return set_global_err(JR_E_UNSUPPORTED, "image sums are built for FP64 pairs only");
  const int ch = pot->eval_chunks;
  if (ch != 0 && ch != 1 && ch != 2 && ch != 4 && ch != 8)
    return set_global_err(JR_E_ARG, "eval_chunks must be 0, 1, 2, 4 or 8, got %d", ch);
  if (ch > 1 && (pot_images(pot) || pot->precision != JR_PREC_FP64))
    return set_global_err(JR_E_UNSUPPORTED,
                          "eval_chunks > 1 is built for FP64 pairs without images");
  return JR_OK;
}

// Arguments of jr_eval_energy_forces and its allocation-free variant.
inline int check_eval_args(const jr_pot* pot, int B, int Nmax, const double* pos,
                           const int32_t* types, const int32_t* n_atoms, const double* cell) {
  int rc = validate_pot(pot);
  if (rc) return rc;
  if (B < 0 || Nmax < 1 || Nmax > 256)
    return set_global_err(JR_E_ARG, "need B >= 0 and 1 <= Nmax <= 256");
  if (B > 0 && (!pos || !types || !n_atoms || !cell))
    return set_global_err(JR_E_ARG, "pos/types/n_atoms/cell must be set");
  for (int b = 0; b < B; ++b) {
    if (n_atoms[b] < 0 || n_atoms[b] > Nmax)
      return set_global_err(JR_E_ARG, "n_atoms[%d]=%d outside [0, %d]", b, n_atoms[b], Nmax);
    for (int i = 0; i < n_atoms[b]; ++i) {
      const int t = types[(size_t)b * Nmax + i];
      if (t < 0 || t >= pot->n_species)
        return set_global_err(JR_E_ARG, "species %d out of range", t);
    }
  }
  return JR_OK;
}

}  // namespace jrh
