// Pair-potential helpers shared by the library's translation units:
// kernel-configuration bits of a jr_pot, its validation, and the argument
// checks of the batched evaluators.
#pragma once

#include "jr_host.h"

namespace jrh {

inline int atoms_per_lane(int nmax) { return nmax <= 64 ? 1 : (nmax <= 128 ? 2 : 4); }

inline bool pot_images(const jr_pot* pot) {
  return pot->images[0] > 0 || pot->images[1] > 0 || pot->images[2] > 0;
}

// kernel configuration bits 0-2 (jr_launch.h)
inline int pot_cfg(const jr_pot* pot) {
  return (pot->n_species > 1 ? 1 : 0) |
         (pot->precision == JR_PREC_FP32_PAIRS ? 2 : 0) |
         (pot_images(pot) ? 4 : 0);
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
