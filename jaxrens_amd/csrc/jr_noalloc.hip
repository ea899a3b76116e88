// jr_eval_energy_forces_noalloc: the batched energy (+forces) evaluation of
// jr_eval_energy_forces without any device allocation.  Its staging buffers
// (inputs copied to the device, outputs before the copy back) live in one
// file-scope device array per GPU, `scratch`, of a fixed size; each device's
// address of it is taken once, with that device current, and kept.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "jr_launch.h"
#include "jr_pot.h"

using namespace jr;
using namespace jrh;

namespace {

constexpr size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// Layout of one call's staging area for B configurations of Nmax slots and
// S species; every part 256-byte aligned.
struct Layout {
  size_t pos, F, cell, pr, mu, U, H, e4, e24, s2, types, n_atoms, total;
  Layout(size_t B, size_t Nmax, size_t S) {
    size_t o = 0;
    auto take = [&](size_t bytes) { const size_t at = o; o += align256(bytes); return at; };
    pos = take(B * Nmax * 3 * sizeof(double));
    F = take(B * Nmax * 3 * sizeof(double));
    cell = take(B * 9 * sizeof(double));
    pr = take(B * sizeof(double));
    mu = take(B * S * sizeof(double));
    U = take(B * sizeof(double));
    H = take(B * sizeof(double));
    e4 = take(S * S * sizeof(double));
    e24 = take(S * S * sizeof(double));
    s2 = take(S * S * sizeof(double));
    types = take(B * Nmax * sizeof(int));
    n_atoms = take(B * sizeof(int));
    total = o;
  }
};

// Sized for the largest call the header allows: B * Nmax atom slots up to
// JR_NOALLOC_MAX_SLOTS and B up to JR_NOALLOC_MAX_CONFIGS with
// JR_MAX_SPECIES species (a layout of fewer configurations never needs more).
constexpr size_t kScratchBytes =
    align256((size_t)JR_NOALLOC_MAX_SLOTS * 3 * sizeof(double)) * 2 +
    align256((size_t)JR_NOALLOC_MAX_CONFIGS * 9 * sizeof(double)) +
    align256((size_t)JR_NOALLOC_MAX_CONFIGS * sizeof(double)) * 3 +
    align256((size_t)JR_NOALLOC_MAX_CONFIGS * JR_MAX_SPECIES * sizeof(double)) +
    align256((size_t)JR_MAX_SPECIES * JR_MAX_SPECIES * sizeof(double)) * 3 +
    align256((size_t)JR_NOALLOC_MAX_SLOTS * sizeof(int)) +
    align256((size_t)JR_NOALLOC_MAX_CONFIGS * sizeof(int));

// One copy per GPU, held until the process ends.  In host code this name is
// only a handle for hipGetSymbolAddress; the device address is scratch_ptr.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// This device's address of `scratch` (taken on its first call, with the
// device current); NULL and the global error set on failure.
unsigned char* scratch_ptr(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_global_err(JR_E_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", device,
                     hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_ptr[device] = static_cast<unsigned char*>(p);
  }
  return g_scratch_ptr[device];
}

}  // namespace

extern "C" int jr_eval_energy_forces_noalloc(const jr_pot* pot, int32_t device, int32_t B,
                                             int32_t Nmax, const double* pos,
                                             const int32_t* types, const int32_t* n_atoms,
                                             const double* cell, const double* pressure,
                                             const double* mu, double* U, double* H,
                                             double* F) {
  int rc = check_eval_args(pot, B, Nmax, pos, types, n_atoms, cell);
  if (rc) return rc;
  if ((size_t)B * (size_t)Nmax > (size_t)JR_NOALLOC_MAX_SLOTS || B > JR_NOALLOC_MAX_CONFIGS)
    return set_global_err(JR_E_ARG,
                          "jr_eval_energy_forces_noalloc: B = %d configurations of Nmax = %d "
                          "exceed the scratch capacity (B * Nmax <= %d, B <= %d)",
                          B, Nmax, JR_NOALLOC_MAX_SLOTS, JR_NOALLOC_MAX_CONFIGS);
  if (device < 0 || device >= kMaxDevices)
    return set_global_err(JR_E_ARG, "device %d outside [0, %d)", device, kMaxDevices);
  if (B == 0) return JR_OK;
  HIPCHK_G(hipSetDevice(device));
  unsigned char* base = scratch_ptr(device);
  if (!base) return JR_E_HIP;
  const int S = pot->n_species;
  const Layout L((size_t)B, (size_t)Nmax, (size_t)S);
  if (L.total > kScratchBytes)
    return set_global_err(JR_E_ARG, "jr_eval_energy_forces_noalloc: %zu staging bytes exceed "
                          "the scratch array's %zu", L.total, kScratchBytes);
  auto at = [&](size_t off) { return base + off; };
  // the default stream: no stream, event or buffer is created
  hipStream_t s = nullptr;
  auto h2d = [&](size_t off, const void* src, size_t bytes) {
    return hipMemcpyAsync(at(off), src, bytes, hipMemcpyHostToDevice, s);
  };
  HIPCHK_G(h2d(L.pos, pos, sizeof(double) * B * Nmax * 3));
  HIPCHK_G(h2d(L.cell, cell, sizeof(double) * B * 9));
  HIPCHK_G(h2d(L.types, types, sizeof(int) * B * Nmax));
  HIPCHK_G(h2d(L.n_atoms, n_atoms, sizeof(int) * B));
  HIPCHK_G(h2d(L.e4, pot->eps4, sizeof(double) * S * S));
  HIPCHK_G(h2d(L.e24, pot->eps24, sizeof(double) * S * S));
  HIPCHK_G(h2d(L.s2, pot->sig2, sizeof(double) * S * S));
  if (pressure) HIPCHK_G(h2d(L.pr, pressure, sizeof(double) * B));
  if (mu) HIPCHK_G(h2d(L.mu, mu, sizeof(double) * B * S));
  EvalArgs a;
  std::memset(&a, 0, sizeof(a));
  a.pos = reinterpret_cast<const double*>(at(L.pos));
  a.types = reinterpret_cast<const int*>(at(L.types));
  a.n_atoms = reinterpret_cast<const int*>(at(L.n_atoms));
  a.cell = reinterpret_cast<const double*>(at(L.cell));
  a.pressure = pressure ? reinterpret_cast<const double*>(at(L.pr)) : nullptr;
  a.mu = mu ? reinterpret_cast<const double*>(at(L.mu)) : nullptr;
  a.U = reinterpret_cast<double*>(at(L.U));
  a.H = reinterpret_cast<double*>(at(L.H));
  a.F = F ? reinterpret_cast<double*>(at(L.F)) : nullptr;
  a.B = B;
  a.Nmax = Nmax;
  a.S = S;
  a.rc2 = pot->rc2;
  a.eps4 = reinterpret_cast<const double*>(at(L.e4));
  a.eps24 = reinterpret_cast<const double*>(at(L.e24));
  a.sig2 = reinterpret_cast<const double*>(at(L.s2));
  set_images(a.img, *pot);
  a.chunks = pot_chunks(pot);
  a.team = team_waves(atoms_per_lane(Nmax), a.chunks, S, device);
  launch_eval(atoms_per_lane(Nmax), pot_cfg(pot), a, s);
  HIPCHK_G(hipGetLastError());
  auto d2h = [&](void* dst, size_t off, size_t bytes) {
    return hipMemcpyAsync(dst, at(off), bytes, hipMemcpyDeviceToHost, s);
  };
  if (U) HIPCHK_G(d2h(U, L.U, sizeof(double) * B));
  if (H) HIPCHK_G(d2h(H, L.H, sizeof(double) * B));
  if (F) HIPCHK_G(d2h(F, L.F, sizeof(double) * B * Nmax * 3));
  HIPCHK_G(hipStreamSynchronize(s));
  return JR_OK;
}
