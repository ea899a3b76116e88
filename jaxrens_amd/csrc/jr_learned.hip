// Host side of the learned part of libjrhip: the message-passing judge
// (batched energy and forces), the surrogate (evaluation, gradient, Adam)
// and the surrogate-screened NS walk of the learned model.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "jr_handle.h"
#include "jr_mpnn.h"

using namespace jr;
using namespace jrh;

namespace {

int validate_mpnn(const jr_mpnn* net) {
  if (!net) return set_global_err(JR_E_ARG, "mpnn is NULL");
  if (net->n_species < 1 || net->n_species > MP_MAXS)
    return set_global_err(JR_E_UNSUPPORTED, "mpnn n_species must be in [1, %d]", MP_MAXS);
  if (net->channels != MP_C || net->n_basis != MP_NB)
    return set_global_err(JR_E_UNSUPPORTED, "mpnn needs %d channels and %d radial functions",
                          MP_C, MP_NB);
  if (!(net->rc > 0.0) || !(net->core_rc > 0.0) || !(net->core_sig > 0.0))
    return set_global_err(JR_E_ARG, "mpnn cutoffs must be positive");
  if (!net->emb || !net->W1 || !net->U1 || !net->c1 || !net->W2 || !net->U2 ||
      !net->c2 || !net->wo)
    return set_global_err(JR_E_ARG, "mpnn weights must not be NULL");
  return JR_OK;
}

// Cubic cell with edge >= rc (the kernel enumerates at most two images per axis).
bool cubic_ok(const double* c, double rc) {
  return c[1] == 0.0 && c[2] == 0.0 && c[3] == 0.0 && c[5] == 0.0 && c[6] == 0.0 &&
         c[7] == 0.0 && c[0] == c[4] && c[0] == c[8] && c[0] >= rc;
}

// Device copy of the weights (one buffer) and the kernel's parameter view.
struct MPNNWeights {
  DBuf<float> buf;
  MPNNDev dev;
  hipError_t upload(const jr_mpnn* net, hipStream_t s) {
    const int S = net->n_species;
    const size_t sz[8] = {(size_t)S * MP_C, (size_t)MP_NB * MP_C, (size_t)MP_C * MP_C,
                          (size_t)MP_C, (size_t)MP_NB * MP_C, (size_t)MP_C * MP_C,
                          (size_t)MP_C, (size_t)MP_C};
    const float* src[8] = {net->emb, net->W1, net->U1, net->c1,
                           net->W2, net->U2, net->c2, net->wo};
    size_t total = 0;
    for (size_t v : sz) total += v;
    hipError_t e = buf.alloc(total);
    if (e != hipSuccess) return e;
    size_t off = 0;  // packed in the order of MPW<S>
    for (int k = 0; k < 8; ++k) {
      e = hipMemcpyAsync(buf.p + off, src[k], sizeof(float) * sz[k],
                         hipMemcpyHostToDevice, s);
      if (e != hipSuccess) return e;
      off += sz[k];
    }
    dev.S = S;
    dev.rc = (float)net->rc;
    dev.core_eps = (float)net->core_eps;
    dev.core_sig = (float)net->core_sig;
    dev.core_rc = (float)net->core_rc;
    dev.pref = (float)std::sqrt(2.0 / net->rc);
    dev.core_sh = (float)std::pow(net->core_sig / net->core_rc, 12.0);
    dev.inv_rc = (float)(1.0 / net->rc);
    dev.pi_rc = (float)(3.14159265358979323846 / net->rc);
    return hipSuccess;
  }
};

template <int S, bool FORCES>
hipError_t launch_mpnn_t(const MPNNArgs& a, const float* w, int grid, hipStream_t s) {
  const size_t lds = mp_lds_bytes(FORCES);
  hipError_t e = hipFuncSetAttribute((const void*)mpnn_kernel<S, FORCES>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((mpnn_kernel<S, FORCES>), dim3(grid), dim3(MP_THREADS), lds, s, a, w);
  return hipGetLastError();
}

// Persistent blocks per CU the judge's scratch lists are sized for: the
// energy-only kernel fits two blocks per CU, the force kernel one.
constexpr int kMpBlocksPerCu = 2;

// Launch the judge on items [0, a.B) with `grid` persistent blocks (at most
// the blocks its neighbour-list scratch holds).
hipError_t launch_mpnn(const MPNNArgs& a, const float* w, int grid, hipStream_t s) {
  if (a.net.S == 1)
    return a.F ? launch_mpnn_t<1, true>(a, w, grid, s) : launch_mpnn_t<1, false>(a, w, grid, s);
  return a.F ? launch_mpnn_t<2, true>(a, w, grid, s) : launch_mpnn_t<2, false>(a, w, grid, s);
}

int device_cus(int device) {
  int n = 0;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    return 256;
  return n > 0 ? n : 256;
}

}  // namespace

extern "C" int jr_mpnn_eval(const jr_mpnn* net, int32_t device, int32_t B, int32_t Nmax,
                            const double* pos, const int32_t* types,
                            const int32_t* n_atoms, const double* cell,
                            const double* pressure, double* U, double* H, double* F) {
  int rc = validate_mpnn(net);
  if (rc) return rc;
  if (B < 0 || Nmax < 1 || Nmax > MP_THREADS)
    return set_global_err(JR_E_ARG, "need B >= 0 and 1 <= Nmax <= %d", MP_THREADS);
  if (!pos || !types || !n_atoms || !cell)
    return set_global_err(JR_E_ARG, "pos/types/n_atoms/cell must be set");
  if (B == 0) return JR_OK;
  for (int b = 0; b < B; ++b) {
    if (n_atoms[b] < 0 || n_atoms[b] > Nmax)
      return set_global_err(JR_E_ARG, "n_atoms[%d]=%d outside [0, %d]", b, n_atoms[b], Nmax);
    if (!cubic_ok(cell + (size_t)b * 9, net->rc))
      return set_global_err(JR_E_UNSUPPORTED,
                            "configuration %d: the message-passing judge needs a cubic "
                            "cell with edge >= rc = %g", b, net->rc);
    for (int i = 0; i < n_atoms[b]; ++i) {
      const int t = types[(size_t)b * Nmax + i];
      if (t < 0 || t >= net->n_species)
        return set_global_err(JR_E_ARG, "species %d out of range", t);
    }
  }
  HIPCHK_G(hipSetDevice(device));
  hipStream_t s;
  HIPCHK_G(hipStreamCreate(&s));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } guard{s};
  MPNNWeights w;
  HIPCHK_G(w.upload(net, s));
  const int grid = std::min(B, (F ? 1 : kMpBlocksPerCu) * device_cus(device));
  DBuf<double> d_pos, d_cell, d_U, d_F;
  DBuf<int> d_t, d_n, d_of;
  DBuf<unsigned> d_nbl;
  HIPCHK_G(d_pos.alloc((size_t)B * Nmax * 3));
  HIPCHK_G(d_cell.alloc((size_t)B * 9));
  HIPCHK_G(d_t.alloc((size_t)B * Nmax));
  HIPCHK_G(d_n.alloc((size_t)B));
  HIPCHK_G(d_U.alloc((size_t)B));
  HIPCHK_G(d_of.alloc(1));
  HIPCHK_G(d_nbl.alloc((size_t)grid * MP_MAXNB * MP_THREADS));
  if (F) HIPCHK_G(d_F.alloc((size_t)B * Nmax * 3));
  HIPCHK_G(hipMemcpyAsync(d_pos.p, pos, sizeof(double) * B * Nmax * 3, hipMemcpyHostToDevice, s));
  HIPCHK_G(hipMemcpyAsync(d_cell.p, cell, sizeof(double) * B * 9, hipMemcpyHostToDevice, s));
  HIPCHK_G(hipMemcpyAsync(d_t.p, types, sizeof(int) * B * Nmax, hipMemcpyHostToDevice, s));
  HIPCHK_G(hipMemcpyAsync(d_n.p, n_atoms, sizeof(int) * B, hipMemcpyHostToDevice, s));
  HIPCHK_G(hipMemsetAsync(d_of.p, 0, sizeof(int), s));
  if (F) HIPCHK_G(hipMemsetAsync(d_F.p, 0, sizeof(double) * B * Nmax * 3, s));
  MPNNArgs a;
  a.net = w.dev;
  a.B = B;
  a.Nmax = Nmax;
  a.pos = d_pos.p;
  a.types = d_t.p;
  a.n_atoms = d_n.p;
  a.cell = d_cell.p;
  a.cfg_index = nullptr;
  a.count = nullptr;
  a.E = d_U.p;
  a.F = F ? d_F.p : nullptr;
  a.nbl = d_nbl.p;
  a.overflow = d_of.p;
  HIPCHK_G(launch_mpnn(a, w.buf.p, grid, s));
  std::vector<double> u(B);
  int of = 0;
  HIPCHK_G(hipMemcpyAsync(u.data(), d_U.p, sizeof(double) * B, hipMemcpyDeviceToHost, s));
  HIPCHK_G(hipMemcpyAsync(&of, d_of.p, sizeof(int), hipMemcpyDeviceToHost, s));
  if (F)
    HIPCHK_G(hipMemcpyAsync(F, d_F.p, sizeof(double) * B * Nmax * 3, hipMemcpyDeviceToHost, s));
  HIPCHK_G(hipStreamSynchronize(s));
  if (of) return set_global_err(JR_E_UNSUPPORTED, "neighbour list overflow (> %d images)", MP_MAXNB);
  for (int b = 0; b < B; ++b) {
    if (U) U[b] = u[b];
    if (H) {
      const double* c = cell + (size_t)b * 9;
      const double m0 = c[4] * c[8] - c[5] * c[7], m1 = c[3] * c[8] - c[5] * c[6],
                   m2 = c[3] * c[7] - c[4] * c[6];
      const double V = std::fabs(c[0] * m0 - c[1] * m1 + c[2] * m2);
      H[b] = u[b] + (pressure ? pressure[b] * V : 0.0);
    }
  }
  return JR_OK;
}

// ---------------------------------------------------------------- surrogate
#include "jr_surrogate.h"

namespace {

// Device parameters of the surrogate: FP32 master copy and BF16 operands.
struct SurParams {
  DBuf<float> P;
  DBuf<__bf16> W1t, W1r, W2t, W2r;
  hipError_t alloc() {
    hipError_t e;
    if ((e = P.alloc(SG_NP))) return e;
    if ((e = W1t.alloc((size_t)SG_H * SG_D))) return e;
    if ((e = W1r.alloc((size_t)SG_DP * SG_H))) return e;
    if ((e = W2t.alloc((size_t)SG_H * SG_H))) return e;
    return W2r.alloc((size_t)SG_H * SG_H);
  }
  void cast(hipStream_t s) {
    hipLaunchKernelGGL(sur_cast_kernel, dim3((SG_H * SG_H + 255) / 256), dim3(256), 0, s,
                       P.p, W1t.p, W1r.p, W2t.p, W2r.p);
  }
  SurDev dev() const { return SurDev{P.p, W1t.p, W1r.p, W2t.p, W2r.p}; }
};

// Baseline pair term (eps, sig, rc) into the kernels' parameters; base ==
// NULL or eps == 0 switches it off.
int set_baseline(SurCfgArgs& a, const double* base) {
  a.b_eps = a.b_sig2 = a.b_rc2 = a.b_shift = 0.0f;
  if (!base || base[0] == 0.0) return JR_OK;
  if (!(base[1] > 0.0) || !(base[2] > 0.0) || base[2] > (double)SG_RD)
    return set_global_err(JR_E_ARG, "baseline needs sig > 0 and 0 < rc <= %g", (double)SG_RD);
  a.b_eps = (float)base[0];
  a.b_sig2 = (float)(base[1] * base[1]);
  a.b_rc2 = (float)(base[2] * base[2]);
  a.b_shift = (float)std::pow(base[1] / base[2], 12.0);
  return JR_OK;
}

// The surrogate's passes over the items of c (c.X, c.Ebase and, with F,
// c.dX set by the caller): E [cfg] and F [cfg][Nmax][3] (NULL: energy only).
hipError_t sur_passes(const SurDev& net, SurCfgArgs c, float* erow, double* E, double* F,
                      hipStream_t s) {
  if (c.B < 1) return hipSuccess;
  hipLaunchKernelGGL(sur_desc_kernel, dim3(c.B), dim3(256), 0, s, c);
  SurMlpArgs m;
  m.net = net;
  m.B = c.B;
  m.Nmax = c.Nmax;
  m.tiles = (c.Nmax + 31) / 32;
  m.cfg_index = c.cfg_index;
  m.count = c.count;
  m.X = c.X;
  m.erow = erow;
  m.dX = F ? const_cast<float*>(c.dX) : nullptr;
  hipLaunchKernelGGL(sur_mlp_kernel, dim3(c.B * m.tiles), dim3(64), 0, s, m);
  hipLaunchKernelGGL(sur_energy_kernel, dim3((c.B + 3) / 4), dim3(256), 0, s, c.B, c.Nmax,
                     c.cfg_index, c.count, c.n_atoms, erow, net.P, c.Ebase, E);
  if (F) {
    c.F = F;
    const size_t lds = sizeof(float) * (4 * 256 + (size_t)c.Nmax * 64);
    hipError_t e = hipFuncSetAttribute((const void*)sur_force_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(sur_force_kernel, dim3(c.B), dim3(256), lds, s, c);
  }
  return hipGetLastError();
}

struct StreamHolder {
  hipStream_t s = nullptr;
  ~StreamHolder() {
    if (s) (void)hipStreamDestroy(s);
  }
};

}  // namespace

extern "C" int jr_sur_eval(const float* params, const double* base, int32_t device,
                           int32_t B, int32_t Nmax, const double* pos,
                           const int32_t* types, const int32_t* n_atoms, const double* cell,
                           int32_t n_species, double* U, double* F, uint16_t* x_out) {
  if (!params || !pos || !types || !n_atoms || !cell)
    return set_global_err(JR_E_ARG, "params/pos/types/n_atoms/cell must be set");
  if (B < 0 || Nmax < 1 || Nmax > 256)
    return set_global_err(JR_E_ARG, "need B >= 0 and 1 <= Nmax <= 256");
  if (n_species < 1 || n_species > 2)
    return set_global_err(JR_E_UNSUPPORTED, "surrogate supports 1 or 2 species");
  if (B == 0) return JR_OK;
  for (int b = 0; b < B; ++b) {
    if (n_atoms[b] < 0 || n_atoms[b] > Nmax)
      return set_global_err(JR_E_ARG, "n_atoms[%d]=%d outside [0, %d]", b, n_atoms[b], Nmax);
    if (!cubic_ok(cell + (size_t)b * 9, (double)SG_RD))
      return set_global_err(JR_E_UNSUPPORTED, "configuration %d: cubic cell with edge >= %g needed",
                            b, (double)SG_RD);
    for (int i = 0; i < n_atoms[b]; ++i) {
      const int t = types[(size_t)b * Nmax + i];
      if (t < 0 || t >= n_species) return set_global_err(JR_E_ARG, "species %d out of range", t);
    }
  }
  SurCfgArgs c;
  std::memset(&c, 0, sizeof(c));
  int brc = set_baseline(c, base);
  if (brc) return brc;
  HIPCHK_G(hipSetDevice(device));
  StreamHolder sh;
  HIPCHK_G(hipStreamCreate(&sh.s));
  hipStream_t s = sh.s;
  SurParams w;
  HIPCHK_G(w.alloc());
  HIPCHK_G(hipMemcpyAsync(w.P.p, params, sizeof(float) * SG_NP, hipMemcpyHostToDevice, s));
  w.cast(s);
  HIPCHK_G(hipGetLastError());
  const size_t rows = (size_t)B * Nmax;
  DBuf<double> d_pos, d_cell, d_U, d_F, d_Eb;
  DBuf<int> d_t, d_n;
  DBuf<__bf16> d_x;
  DBuf<float> d_erow, d_dX;
  HIPCHK_G(d_pos.alloc(rows * 3));
  HIPCHK_G(d_cell.alloc((size_t)B * 9));
  HIPCHK_G(d_t.alloc(rows));
  HIPCHK_G(d_n.alloc((size_t)B));
  HIPCHK_G(d_U.alloc((size_t)B));
  HIPCHK_G(d_Eb.alloc((size_t)B));
  HIPCHK_G(d_x.alloc(rows * SG_D));
  HIPCHK_G(d_erow.alloc(rows));
  if (F) {
    HIPCHK_G(d_F.alloc(rows * 3));
    HIPCHK_G(d_dX.alloc(rows * 64));
  }
  HIPCHK_G(hipMemcpyAsync(d_pos.p, pos, sizeof(double) * rows * 3, hipMemcpyHostToDevice, s));
  HIPCHK_G(hipMemcpyAsync(d_cell.p, cell, sizeof(double) * B * 9, hipMemcpyHostToDevice, s));
  HIPCHK_G(hipMemcpyAsync(d_t.p, types, sizeof(int) * rows, hipMemcpyHostToDevice, s));
  HIPCHK_G(hipMemcpyAsync(d_n.p, n_atoms, sizeof(int) * B, hipMemcpyHostToDevice, s));
  c.B = B;
  c.Nmax = Nmax;
  c.pos = d_pos.p;
  c.types = d_t.p;
  c.n_atoms = d_n.p;
  c.cell = d_cell.p;
  c.X = d_x.p;
  c.dX = d_dX.p;
  c.Ebase = d_Eb.p;
  HIPCHK_G(sur_passes(w.dev(), c, d_erow.p, d_U.p, F ? d_F.p : nullptr, s));
  if (U) HIPCHK_G(hipMemcpyAsync(U, d_U.p, sizeof(double) * B, hipMemcpyDeviceToHost, s));
  if (F) HIPCHK_G(hipMemcpyAsync(F, d_F.p, sizeof(double) * rows * 3, hipMemcpyDeviceToHost, s));
  if (x_out)
    HIPCHK_G(hipMemcpyAsync(x_out, d_x.p, sizeof(__bf16) * rows * SG_D, hipMemcpyDeviceToHost, s));
  HIPCHK_G(hipStreamSynchronize(s));
  return JR_OK;
}

extern "C" int jr_sur_grad(const float* params, int32_t device, int32_t n_rows,
                           const uint16_t* x, const float* c, float* e_out, float* grad_out) {
  if (!params || !x || n_rows < 1) return set_global_err(JR_E_ARG, "need params, x and n_rows >= 1");
  if (grad_out && !c) return set_global_err(JR_E_ARG, "grad_out needs c");
  HIPCHK_G(hipSetDevice(device));
  StreamHolder sh;
  HIPCHK_G(hipStreamCreate(&sh.s));
  hipStream_t s = sh.s;
  SurParams w;
  HIPCHK_G(w.alloc());
  HIPCHK_G(hipMemcpyAsync(w.P.p, params, sizeof(float) * SG_NP, hipMemcpyHostToDevice, s));
  w.cast(s);
  const int n_chunks = (n_rows + 63) / 64;
  DBuf<__bf16> d_x;
  DBuf<float> d_e, d_c, d_part, d_g;
  HIPCHK_G(d_x.alloc((size_t)n_rows * SG_D));
  HIPCHK_G(d_e.alloc((size_t)n_rows));
  HIPCHK_G(hipMemcpyAsync(d_x.p, x, sizeof(uint16_t) * n_rows * SG_D, hipMemcpyHostToDevice, s));
  SurRowsArgs a;
  a.net = w.dev();
  a.n_rows = n_rows;
  a.x = d_x.p;
  a.e = d_e.p;
  a.c = nullptr;
  a.gpart = nullptr;
  const int grid = std::min(n_chunks, 4 * device_cus(device));
  hipLaunchKernelGGL(sur_rows_forward_kernel, dim3(grid), dim3(64), 0, s, a);
  HIPCHK_G(hipGetLastError());
  if (grad_out) {
    HIPCHK_G(d_c.alloc((size_t)n_rows));
    HIPCHK_G(d_part.alloc((size_t)n_chunks * SG_NPP));
    HIPCHK_G(d_g.alloc(SG_NP));
    HIPCHK_G(hipMemcpyAsync(d_c.p, c, sizeof(float) * n_rows, hipMemcpyHostToDevice, s));
    a.c = d_c.p;
    a.gpart = d_part.p;
    hipLaunchKernelGGL(sur_rows_backward_kernel, dim3(n_chunks), dim3(64), 0, s, a);
    HIPCHK_G(hipGetLastError());
    hipLaunchKernelGGL(sur_reduce_kernel, dim3((SG_NP + 255) / 256), dim3(256), 0, s,
                       d_part.p, n_chunks, d_g.p);
    HIPCHK_G(hipGetLastError());
    HIPCHK_G(hipMemcpyAsync(grad_out, d_g.p, sizeof(float) * SG_NP, hipMemcpyDeviceToHost, s));
  }
  if (e_out) HIPCHK_G(hipMemcpyAsync(e_out, d_e.p, sizeof(float) * n_rows, hipMemcpyDeviceToHost, s));
  HIPCHK_G(hipStreamSynchronize(s));
  return JR_OK;
}

extern "C" int jr_sur_adam(int32_t device, float* params, float* m, float* v,
                           const float* grad, double lr, int32_t t) {
  if (!params || !m || !v || !grad || t < 1) return set_global_err(JR_E_ARG, "bad Adam arguments");
  HIPCHK_G(hipSetDevice(device));
  StreamHolder sh;
  HIPCHK_G(hipStreamCreate(&sh.s));
  hipStream_t s = sh.s;
  DBuf<float> d_p, d_m, d_v, d_g;
  HIPCHK_G(d_p.alloc(SG_NP));
  HIPCHK_G(d_m.alloc(SG_NP));
  HIPCHK_G(d_v.alloc(SG_NP));
  HIPCHK_G(d_g.alloc(SG_NP));
  HIPCHK_G(hipMemcpyAsync(d_p.p, params, sizeof(float) * SG_NP, hipMemcpyHostToDevice, s));
  HIPCHK_G(hipMemcpyAsync(d_m.p, m, sizeof(float) * SG_NP, hipMemcpyHostToDevice, s));
  HIPCHK_G(hipMemcpyAsync(d_v.p, v, sizeof(float) * SG_NP, hipMemcpyHostToDevice, s));
  HIPCHK_G(hipMemcpyAsync(d_g.p, grad, sizeof(float) * SG_NP, hipMemcpyHostToDevice, s));
  const float c1 = (float)(1.0 - std::pow(0.9, t)), c2 = (float)(1.0 - std::pow(0.999, t));
  hipLaunchKernelGGL(sur_adam_kernel, dim3((SG_NP + 255) / 256), dim3(256), 0, s, d_p.p, d_m.p,
                     d_v.p, d_g.p, (float)lr, c1, c2);
  HIPCHK_G(hipGetLastError());
  HIPCHK_G(hipMemcpyAsync(params, d_p.p, sizeof(float) * SG_NP, hipMemcpyDeviceToHost, s));
  HIPCHK_G(hipMemcpyAsync(m, d_m.p, sizeof(float) * SG_NP, hipMemcpyDeviceToHost, s));
  HIPCHK_G(hipMemcpyAsync(v, d_v.p, sizeof(float) * SG_NP, hipMemcpyDeviceToHost, s));
  HIPCHK_G(hipStreamSynchronize(s));
  return JR_OK;
}

// ------------------------------------------------------- learned-model walk
#include "jr_lwalk.h"

namespace jrl {

struct LearnedState {
  jr_learned cfg{};
  int mode = JR_LEARNED_PLAIN;
  bool lj = false;  // judge: the handle's pair potential (else the message-passing model)
  int S = 1;        // species
  std::vector<jr_move> moves;
  MPNNWeights judge;
  SurParams sur;
  DBuf<float> adam_m, adam_v;
  long long adam_t = 0;
  long long train_counter = 0;
  int C = 0, Nmax = 0, n_cu = 0;
  // chain work arrays
  DBuf<double> cpos, cdir, x0, d0, npos, F, mom, ccell, ncell, cH, cHs, cem, cpres, thr_steer,
      thr_screen, U, Us, Ub, Hy, Hsy, K0;
  DBuf<int> ctypes, cn, code, code2, accept, reason, idx, idx2, cnt;
  DBuf<__bf16> xstage;     // [C][Nmax][80] descriptor rows of the latest surrogate pass
  DBuf<float> erow, dXg;    // [C][Nmax] row energies, [C][Nmax][64] dE/dG rows
  DBuf<unsigned> nbl;
  DBuf<int> overflow;
  DBuf<double> dcontour;  // bound on U / N at this iteration's contours
  double vol_lo = 0.0, vol_hi = 0.0;  // volume-per-atom range of the volume moves (0: none)
  // training ring and minibatch
  DBuf<__bf16> ring;
  DBuf<double> ringE;
  DBuf<int> ringN;
  DBuf<RingState> rs;
  DBuf<int> tidx, tn;
  // trial moves: sampled walkers, their draws and the step sizes with the
  // trial move's column replaced (allocated on the first trial)
  DBuf<int> trial_idx;
  DBuf<double> trial_norm, trial_unif, trial_steps;
  std::vector<int> trial_host;
  DBuf<__bf16> tX;
  DBuf<double> tE, tsq;
  DBuf<float> te, tc, gpart, grad;
  // statistics
  DBuf<LWStats> dstats;
  DBuf<double> hs;            // [R][LW_HS] held-out sums of the running iteration
  DBuf<double> dbias, dsig;   // [R] residual bias / spread per atom, on the device
  std::vector<double> bias, sig;
  double ca = 0.0, cv = 0.0;  // fitted error model e = ca + cv V / N
  jr_learned_stats acc{};
  double rmse = INFINITY;     // held-out per-atom RMSE after the error model
  long long rmse_n = 0;
  long long stored = 0;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  std::vector<int> sched;

  ~LearnedState() {
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};

namespace {

unsigned blocks_for(size_t n, unsigned per = 256) {
  return (unsigned)std::max<size_t>(1, (n + per - 1) / per);
}

hipError_t zero_stats(LearnedState* L, hipStream_t s) {
  hipError_t e = hipMemsetAsync(L->dstats.p, 0, sizeof(LWStats), s);
  if (e != hipSuccess) return e;
  e = hipMemsetAsync(L->hs.p, 0, sizeof(double) * L->hs.n, s);
  if (e != hipSuccess) return e;
  return hipMemsetAsync(L->overflow.p, 0, sizeof(int), s);
}

MPNNArgs judge_args(LearnedState* L, const double* pos, const double* cell, int B,
                    const int* types, const int* n, const int* idx, const int* count,
                    double* E, double* F) {
  MPNNArgs a;
  a.net = L->judge.dev;
  a.B = B;
  a.Nmax = L->Nmax;
  a.pos = pos;
  a.types = types;
  a.n_atoms = n;
  a.cell = cell;
  a.cfg_index = idx;
  a.count = count;
  a.E = E;
  a.F = F;
  a.nbl = L->nbl.p;
  a.overflow = L->overflow.p;
  return a;
}

hipError_t run_judge(LearnedState* L, const MPNNArgs& a, hipStream_t s) {
  const int per_cu = a.F ? 1 : kMpBlocksPerCu;
  return launch_mpnn(a, L->judge.buf.p, std::max(1, std::min(a.B, per_cu * L->n_cu)), s);
}

// The judge's model energies U (and forces F unless NULL) of the listed
// configurations (idx NULL: all B; count NULL: B), indexed by configuration.
hipError_t judge_eval(jr_handle* h, LearnedState* L, const double* pos, const double* cell,
                      const int* types, const int* n, const int* idx, const int* count, int B,
                      double* U, double* F, hipStream_t s) {
  if (L->lj) return lj_eval(h, pos, types, n, cell, idx, count, B, U, F, s);
  return run_judge(L, judge_args(L, pos, cell, B, types, n, idx, count, U, F), s);
}

// The surrogate on the first B chains of the chain arrays (items idx/count,
// or all B): energies E and forces F (NULL: energy only); the descriptor rows
// land in xstage and the baseline energies in Ub for the training ring.
hipError_t run_sur(LearnedState* L, int B, const double* pos, const double* cell,
                   const int* idx, const int* count, double* E, double* F, hipStream_t s) {
  SurCfgArgs c;
  std::memset(&c, 0, sizeof(c));
  const double base[3] = {L->cfg.base_eps, L->cfg.base_sig, L->cfg.base_rc};
  (void)set_baseline(c, base);  // validated at creation
  c.B = B;
  c.Nmax = L->Nmax;
  c.pos = pos;
  c.types = L->ctypes.p;
  c.n_atoms = L->cn.p;
  c.cell = cell;
  c.cfg_index = idx;
  c.count = count;
  c.X = L->xstage.p;
  c.dX = L->dXg.p;
  c.Ebase = L->Ub.p;
  return sur_passes(L->sur.dev(), c, L->erow.p, E, F, s);
}

}  // namespace

int learned_create(jr_handle* h, const jr_config* c) {
  const jr_mpnn* net = c->mpnn;
  if (net) {
    int rc = validate_mpnn(net);
    if (rc) {
      h->err = g_err;
      return rc;
    }
  }
  if (c->n_atoms_max > MP_THREADS)
    return h->fail(JR_E_UNSUPPORTED, "the learned walk takes at most %d atoms", MP_THREADS);
  for (int m = 0; m < c->n_moves; ++m) {
    const int t = c->moves[m].type;
    if (t != JR_MOVE_GMC && t != JR_MOVE_VOLUME && t != JR_MOVE_HMC)
      return h->fail(JR_E_UNSUPPORTED,
                     "move %d: the learned walk runs gmc, volume and hmc moves", m);
    if (t == JR_MOVE_HMC && c->moves[m].n_leapfrog < 1)
      return h->fail(JR_E_ARG, "hmc needs n_leapfrog >= 1");
  }
  jr_learned lc{};
  lc.mode = JR_LEARNED_PLAIN;
  if (c->learned) lc = *c->learned;
  if (lc.mode < JR_LEARNED_PLAIN || lc.mode > JR_LEARNED_SCREENED)
    return h->fail(JR_E_ARG, "learned.mode %d unknown", lc.mode);
  if (lc.mode != JR_LEARNED_PLAIN) {
    if (!lc.sur_params) return h->fail(JR_E_ARG, "learned.sur_params must be set");
    if (lc.ring_capacity < 1 || lc.train_batch < 1 || lc.train_steps < 0 || lc.min_train < 1 ||
        lc.min_heldout < 1 || !(lc.lr > 0.0) || !(lc.rmse_bound >= 0.0) ||
        !(lc.delta_k >= 0.0) || std::isnan(lc.train_max_epa))
      return h->fail(JR_E_ARG, "learned: invalid training settings");
    if (lc.base_eps != 0.0 &&
        (!(lc.base_sig > 0.0) || !(lc.base_rc > 0.0) || lc.base_rc > (double)SG_RD))
      return h->fail(JR_E_ARG, "learned: baseline needs sig > 0 and 0 < rc <= %g",
                     (double)SG_RD);
  }
  const int S = net ? net->n_species : c->pot.n_species;
  if (lc.mode != JR_LEARNED_PLAIN && S > 2)
    return h->fail(JR_E_UNSUPPORTED, "the surrogate takes 1 or 2 species");
  auto* L = new LearnedState();
  h->learned = L;
  L->lj = net == nullptr;
  L->S = S;
  L->cfg = lc;
  L->cfg.sur_params = nullptr;
  L->mode = lc.mode;
  L->moves.assign(c->moves, c->moves + c->n_moves);
  L->Nmax = c->n_atoms_max;
  L->C = c->n_replicas * (1 + c->n_extra);
  L->n_cu = device_cus(c->device);
  L->sched.assign(c->n_mcmc_steps, 0);
  for (const jr_move& mv : L->moves)
    if (mv.type == JR_MOVE_VOLUME) {
      L->vol_lo = L->vol_lo > 0.0 ? std::min(L->vol_lo, mv.min_vpa) : mv.min_vpa;
      L->vol_hi = std::max(L->vol_hi, mv.max_vpa);
    }
  if (net) {
    h->mpnn_cfg = *net;
    h->mpnn_cfg.emb = h->mpnn_cfg.W1 = h->mpnn_cfg.U1 = h->mpnn_cfg.c1 = nullptr;
    h->mpnn_cfg.W2 = h->mpnn_cfg.U2 = h->mpnn_cfg.c2 = h->mpnn_cfg.wo = nullptr;
  }
  h->learned_cfg = L->cfg;
  hipStream_t s = h->stream;
  const size_t C = L->C, Nm = L->Nmax, n3 = C * Nm * 3;
  if (net) HIPCHK(h, L->judge.upload(net, s));
  for (DBuf<double>* b : {&L->cpos, &L->cdir, &L->x0, &L->d0, &L->npos, &L->F, &L->mom})
    HIPCHK(h, b->alloc(n3));
  HIPCHK(h, L->ccell.alloc(C * 9));
  HIPCHK(h, L->ncell.alloc(C * 9));
  for (DBuf<double>* b : {&L->cH, &L->cHs, &L->cem, &L->cpres, &L->thr_steer, &L->thr_screen,
                          &L->U, &L->Us, &L->Ub, &L->Hy, &L->Hsy, &L->K0})
    HIPCHK(h, b->alloc(C));
  const int R = c->n_replicas;
  HIPCHK(h, L->hs.alloc((size_t)R * LW_HS));
  HIPCHK(h, L->dbias.alloc(R));
  HIPCHK(h, L->dsig.alloc(R));
  L->bias.assign(R, 0.0);
  L->sig.assign(R, 0.0);
  HIPCHK(h, L->ctypes.alloc(C * Nm));
  for (DBuf<int>* b : {&L->cn, &L->code, &L->code2, &L->accept, &L->reason, &L->idx, &L->idx2})
    HIPCHK(h, b->alloc(C));
  HIPCHK(h, L->cnt.alloc(4));
  // judge launches use at most kMpBlocksPerCu * n_cu persistent blocks,
  // each with its own neighbour list
  if (net)
    HIPCHK(h, L->nbl.alloc((size_t)kMpBlocksPerCu * L->n_cu * MP_MAXNB * MP_THREADS));
  HIPCHK(h, L->overflow.alloc(1));
  HIPCHK(h, L->dcontour.alloc(1));
  HIPCHK(h, L->dstats.alloc(1));
  HIPCHK(h, L->rs.alloc(1));
  for (auto& e : L->ev) HIPCHK(h, hipEventCreate(&e));
  if (L->mode != JR_LEARNED_PLAIN) {
    HIPCHK(h, L->sur.alloc());
    HIPCHK(h, hipMemcpyAsync(L->sur.P.p, c->learned->sur_params, sizeof(float) * SG_NP,
                             hipMemcpyHostToDevice, s));
    L->sur.cast(s);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, L->adam_m.alloc(SG_NP));
    HIPCHK(h, L->adam_v.alloc(SG_NP));
    HIPCHK(h, hipMemsetAsync(L->adam_m.p, 0, sizeof(float) * SG_NP, s));
    HIPCHK(h, hipMemsetAsync(L->adam_v.p, 0, sizeof(float) * SG_NP, s));
    HIPCHK(h, L->xstage.alloc(C * Nm * SG_D));
    HIPCHK(h, L->erow.alloc(C * Nm));
    HIPCHK(h, L->dXg.alloc(C * Nm * 64));
    HIPCHK(h, hipMemsetAsync(L->xstage.p, 0, sizeof(__bf16) * C * Nm * SG_D, s));
    const size_t cap = (size_t)lc.ring_capacity;
    HIPCHK(h, L->ring.alloc(cap * Nm * SG_D));
    HIPCHK(h, L->ringE.alloc(cap));
    HIPCHK(h, L->ringN.alloc(cap));
    const size_t Bt = (size_t)lc.train_batch, rows = Bt * Nm;
    HIPCHK(h, L->tidx.alloc(Bt));
    HIPCHK(h, L->tn.alloc(Bt));
    HIPCHK(h, L->tE.alloc(Bt));
    HIPCHK(h, L->tsq.alloc(Bt));
    HIPCHK(h, L->tX.alloc(rows * SG_D));
    HIPCHK(h, L->te.alloc(rows));
    HIPCHK(h, L->tc.alloc(rows));
    HIPCHK(h, L->gpart.alloc(((rows + 63) / 64) * SG_NPP));
    HIPCHK(h, L->grad.alloc(SG_NP));
  }
  RingState r0{0, 0, std::max<long long>(1, lc.ring_capacity)};
  HIPCHK(h, hipMemcpyAsync(L->rs.p, &r0, sizeof(r0), hipMemcpyHostToDevice, s));
  HIPCHK(h, zero_stats(L, s));
  HIPCHK(h, hipStreamSynchronize(s));
  return JR_OK;
}

void learned_destroy(jr_handle* h) {
  delete h->learned;
  h->learned = nullptr;
}

int learned_check_population(jr_handle* h, const int32_t* types, const int32_t* n_atoms,
                             const double* cell) {
  const auto& c = h->cfg;
  LearnedState* L = h->learned;
  const size_t W = (size_t)c.n_replicas * c.n_walkers;
  // the message-passing judge needs edge >= rc; the surrogate needs edge >= 3
  const double rc = L->lj ? (L->mode != JR_LEARNED_PLAIN ? (double)SG_RD : 0.0)
                          : h->mpnn_cfg.rc;
  for (size_t w = 0; w < W; ++w) {
    if (!cubic_ok(cell + w * 9, rc))
      return h->fail(JR_E_UNSUPPORTED,
                     "walker %zu: the learned walk needs a cubic cell with edge >= %g", w, rc);
    for (int i = 0; i < n_atoms[w]; ++i)
      if (types[w * c.n_atoms_max + i] >= L->S)
        return h->fail(JR_E_ARG, "species out of range for the judge");
    for (const jr_move& mv : L->moves) {
      if (mv.type != JR_MOVE_VOLUME) continue;
      if (L->lj) {
        if (mv.min_vpa * n_atoms[w] < rc * rc * rc)
          return h->fail(JR_E_UNSUPPORTED,
                         "volume move: min_vpa * N = %g must be >= %g for the surrogate",
                         mv.min_vpa * n_atoms[w], rc * rc * rc);
        continue;
      }
      // every volume that passes the cell check keeps the edge >= rc, and the
      // densest allowed cell keeps the judge's neighbour lists in capacity
      if (mv.min_vpa * n_atoms[w] < rc * rc * rc)
        return h->fail(JR_E_UNSUPPORTED,
                       "volume move: min_vpa * N = %g must be >= rc^3 = %g for the judge",
                       mv.min_vpa * n_atoms[w], rc * rc * rc);
      if (4.18879020478639 * rc * rc * rc / mv.min_vpa > 0.9 * MP_MAXNB)
        return h->fail(JR_E_UNSUPPORTED,
                       "volume move: min_vpa %g allows densities beyond the judge's "
                       "neighbour capacity (%d)", mv.min_vpa, MP_MAXNB);
    }
  }
  return JR_OK;
}

int learned_recompute_energies(jr_handle* h) {
  LearnedState* L = h->learned;
  const auto& c = h->cfg;
  const int W = c.n_replicas * c.n_walkers;
  hipStream_t s = h->stream;
  std::vector<double> pr(W);
  for (int w = 0; w < W; ++w) pr[w] = h->h_pressure[w / c.n_walkers];
  DBuf<double> dU, dP;
  HIPCHK(h, dU.alloc(W));
  HIPCHK(h, dP.alloc(W));
  HIPCHK(h, hipMemcpyAsync(dP.p, pr.data(), sizeof(double) * W, hipMemcpyHostToDevice, s));
  HIPCHK(h, hipMemsetAsync(L->overflow.p, 0, sizeof(int), s));
  HIPCHK(h, judge_eval(h, L, h->pos.p, h->cell.p, h->types.p, h->n_atoms.p, nullptr, nullptr, W,
                       dU.p, nullptr, s));
  hipLaunchKernelGGL(lw_ham_kernel, dim3(blocks_for(W)), dim3(256), 0, s, W, nullptr, nullptr,
                     dU.p, h->cell.p, dP.p, nullptr, 0.0, 0.0, h->energy.p);
  HIPCHK(h, hipGetLastError());
  int of = 0;
  HIPCHK(h, hipMemcpyAsync(&of, L->overflow.p, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(h, hipStreamSynchronize(s));
  if (of) return h->fail(JR_E_UNSUPPORTED, "judge neighbour list overflow (> %d)", MP_MAXNB);
  return JR_OK;
}

namespace {

struct WalkCtx {
  jr_handle* h;
  LearnedState* L;
  hipStream_t s;
  LWChains ch;
  int C, Nm, n_steps;
  bool use_sur;
  double cap;
  // draws [C][n_steps][3][Nm] and [C][n_steps][4]; step sizes [R][n_moves]
  const double* normals;
  const double* uniforms;
  const double* step_sizes;
  // move counters [R][n_moves][JR_NCOUNT] (NULL: none); the walk statistics;
  // training records and the surrogate predictions that only feed them
  int* counters = nullptr;
  bool stats = true, record = true;

  // H of the judge's energies U (all chains, or the listed ones)
  void ham_exact(const int* idx, const int* cnt, const double* U, const double* cell,
                 double* H) const {
    hipLaunchKernelGGL(lw_ham_kernel, dim3(blocks_for(C)), dim3(256), 0, s, C, idx, cnt, U,
                       cell, L->cpres.p, nullptr, 0.0, 0.0, H);
  }
  // corrected H of the surrogate's energies
  void ham_sur(const int* idx, const int* cnt, const double* Us, const double* cell,
               double* H) const {
    hipLaunchKernelGGL(lw_ham_kernel, dim3(blocks_for(C)), dim3(256), 0, s, C, idx, cnt, Us,
                       cell, L->cpres.p, L->cn.p, L->ca, L->cv, H);
  }
};

// Held-out error and ring append over the chains listed in (idx2, cnt[1]);
// cell holds the evaluated configurations' cells.
void record_exact(WalkCtx& w, const int* code, const double* cell) {
  if (!w.record) return;
  LearnedState* L = w.L;
  hipStream_t s = w.s;
  hipLaunchKernelGGL(lw_compact_kernel, dim3(1), dim3(1024), 0, s, w.C, code, 1, 1, L->U.p,
                     L->cn.p, w.cap, L->idx2.p, L->cnt.p + 1);
  hipLaunchKernelGGL(lw_heldout_kernel, dim3(w.h->cfg.n_replicas), dim3(256), 0, s, L->idx2.p,
                     L->cnt.p + 1, w.h->n_walk, L->U.p, L->Us.p, L->cn.p, cell, L->hs.p);
  hipLaunchKernelGGL(lw_ring_append_kernel, dim3(w.C), dim3(256), 0, s, L->idx2.p,
                     L->cnt.p + 1, w.Nm, L->xstage.p, L->U.p, L->Ub.p, L->cn.p, L->rs.p, L->ring.p,
                     L->ringE.p, L->ringN.p);
  hipLaunchKernelGGL(lw_ring_commit_kernel, dim3(1), dim3(64), 0, s, L->cnt.p + 1, L->rs.p);
}

void count_step(WalkCtx& w, int m, const int* code, int eval_const, int eval_code1,
                int grad_const, int screened) {
  if (!w.counters) return;
  LWCountArgs a;
  a.C = w.C;
  a.n_walk = w.ch.n_walk;
  a.n_moves = w.h->cfg.n_moves;
  a.m = m;
  a.reason = w.L->reason.p;
  a.code = code;
  a.eval_const = eval_const;
  a.eval_code1 = eval_code1;
  a.grad_const = grad_const;
  a.counters = w.counters;
  a.stats = w.stats ? w.L->dstats.p : nullptr;
  a.screened = screened;
  hipLaunchKernelGGL(lw_count_kernel, dim3(blocks_for(w.C)), dim3(256), 0, w.s, a);
}

void stat_add(WalkCtx& w, unsigned long long* dst, const int* count, long long n,
              long long mult) {
  if (!w.stats) return;
  hipLaunchKernelGGL(lw_stat_add_kernel, dim3(1), dim3(64), 0, w.s, dst, count, n, mult);
}

hipError_t gmc_step(WalkCtx& w, int t, int m, const jr_move& mv, bool da) {
  LearnedState* L = w.L;
  jr_handle* h = w.h;
  hipStream_t s = w.s;
  const int C = w.C, Nm = w.Nm;
  const size_t xs = (size_t)w.n_steps * 3 * Nm;
  hipLaunchKernelGGL(lw_gmc_begin_kernel, dim3(blocks_for(C, 4)), dim3(256), 0, s, w.ch,
                     L->cdir.p, w.normals + (size_t)t * 3 * Nm, xs, mv.perturb_angle,
                     L->cpos.p, L->x0.p, L->d0.p);
  const unsigned gc = blocks_for((size_t)C * Nm * 3);
  const int nr = mv.n_reflect;
  hipError_t e;
  for (int k = 0; k < nr; ++k) {
    const bool last = k == nr - 1;
    hipLaunchKernelGGL(lw_advance_kernel, dim3(gc), dim3(256), 0, s, w.ch, L->cpos.p,
                       L->cdir.p, w.step_sizes, m);
    if (da) {
      e = run_sur(L, C, L->cpos.p, L->ccell.p, nullptr, nullptr, L->Us.p, L->F.p, s);
      if (e != hipSuccess) return e;
      w.ham_sur(nullptr, nullptr, L->Us.p, L->ccell.p, L->Hsy.p);
      if (!last)
        hipLaunchKernelGGL(lw_reflect_kernel, dim3(blocks_for(C, 4)), dim3(256), 0, s, w.ch,
                           L->cdir.p, L->F.p, L->Hsy.p, L->thr_steer.p);
    } else {
      e = judge_eval(h, L, L->cpos.p, L->ccell.p, L->ctypes.p, L->cn.p, nullptr, nullptr, C,
                     L->U.p, L->F.p, s);
      if (e != hipSuccess) return e;
      w.ham_exact(nullptr, nullptr, L->U.p, L->ccell.p, L->Hy.p);
      if (!last)
        hipLaunchKernelGGL(lw_reflect_kernel, dim3(blocks_for(C, 4)), dim3(256), 0, s, w.ch,
                           L->cdir.p, L->F.p, L->Hy.p, L->cem.p);
    }
  }
  const unsigned gb = blocks_for(C);
  if (da) {
    // stage 1 on the surrogate's verdicts, stage 2 on the judge's energy
    hipLaunchKernelGGL(lw_stage1_kernel, dim3(gb), dim3(256), 0, s, C, nullptr, L->cHs.p,
                       L->Hsy.p, L->thr_screen.p, L->code.p);
    hipLaunchKernelGGL(lw_compact_kernel, dim3(1), dim3(1024), 0, s, C, L->code.p, 1, 0,
                       nullptr, nullptr, 0.0, L->idx.p, L->cnt.p);
    e = judge_eval(h, L, L->cpos.p, L->ccell.p, L->ctypes.p, L->cn.p, L->idx.p, L->cnt.p, C,
                   L->U.p, nullptr, s);
    if (e != hipSuccess) return e;
    w.ham_exact(L->idx.p, L->cnt.p, L->U.p, L->ccell.p, L->Hy.p);
    hipLaunchKernelGGL(lw_accept_kernel, dim3(gb), dim3(256), 0, s, C, L->code.p, L->Hy.p,
                       L->cem.p, L->accept.p, L->reason.p, L->cH.p, L->cHs.p, L->Hsy.p);
    record_exact(w, L->code.p, L->ccell.p);
    count_step(w, m, L->code.p, 0, 1, 0, 1);
    stat_add(w, &L->dstats.p->exact_e, L->cnt.p, 0, 1);
    stat_add(w, &L->dstats.p->sur_e, nullptr, C, nr);
  } else {
    hipLaunchKernelGGL(lw_accept_kernel, dim3(gb), dim3(256), 0, s, C, nullptr, L->Hy.p,
                       L->cem.p, L->accept.p, L->reason.p, L->cH.p, nullptr, nullptr);
    if (w.use_sur && w.record) {
      // the surrogate's prediction at the judged end point: held-out error
      e = run_sur(L, C, L->cpos.p, L->ccell.p, nullptr, nullptr, L->Us.p, nullptr, s);
      if (e != hipSuccess) return e;
      record_exact(w, nullptr, L->ccell.p);
      stat_add(w, &L->dstats.p->sur_e, nullptr, C, 1);
    }
    count_step(w, m, nullptr, nr, 0, nr, 0);
    stat_add(w, &L->dstats.p->exact_fe, nullptr, C, nr);
  }
  hipLaunchKernelGGL(lw_gmc_finish_kernel, dim3(gc), dim3(256), 0, s, w.ch, L->cpos.p,
                     L->cdir.p, L->x0.p, L->d0.p, L->accept.p);
  return hipGetLastError();
}

hipError_t volume_step(WalkCtx& w, int t, int m, const jr_move& mv, bool da) {
  LearnedState* L = w.L;
  jr_handle* h = w.h;
  hipStream_t s = w.s;
  const int C = w.C, Nm = w.Nm;
  LWVolArgs va;
  va.ch = w.ch;
  va.pos = L->cpos.p;
  va.cell = L->ccell.p;
  va.step_sizes = w.step_sizes;
  va.m = m;
  va.xi = w.normals + (size_t)t * 3 * Nm;
  va.xi_stride = (size_t)w.n_steps * 3 * Nm;
  va.u = w.uniforms + (size_t)t * 4;
  va.u_stride = (size_t)w.n_steps * 4;
  va.npos = L->npos.p;
  va.ncell = L->ncell.p;
  va.code = L->code.p;
  va.max_vpa = mv.max_vpa;
  va.min_vpa = mv.min_vpa;
  va.min_aspect = mv.min_aspect;
  va.flat = mv.flat_v_prior;
  const unsigned gb = blocks_for(C);
  hipLaunchKernelGGL(lw_volume_propose_kernel, dim3(blocks_for(C, 4)), dim3(256), 0, s, va);
  hipLaunchKernelGGL(lw_compact_kernel, dim3(1), dim3(1024), 0, s, C, L->code.p, 1, 0, nullptr,
                     nullptr, 0.0, L->idx.p, L->cnt.p);
  hipError_t e;
  const int* fin = L->code.p;  // codes of the final test
  if (da || (w.use_sur && w.record)) {
    e = run_sur(L, C, L->npos.p, L->ncell.p, L->idx.p, L->cnt.p, L->Us.p, nullptr, s);
    if (e != hipSuccess) return e;
    stat_add(w, &L->dstats.p->sur_e, L->cnt.p, 0, 1);
  }
  if (da) {
    w.ham_sur(L->idx.p, L->cnt.p, L->Us.p, L->ncell.p, L->Hsy.p);
    hipLaunchKernelGGL(lw_stage1_kernel, dim3(gb), dim3(256), 0, s, C, L->code.p, L->cHs.p,
                       L->Hsy.p, L->thr_screen.p, L->code2.p);
    hipLaunchKernelGGL(lw_compact_kernel, dim3(1), dim3(1024), 0, s, C, L->code2.p, 1, 0,
                       nullptr, nullptr, 0.0, L->idx.p, L->cnt.p);
    fin = L->code2.p;
  }
  e = judge_eval(h, L, L->npos.p, L->ncell.p, L->ctypes.p, L->cn.p, L->idx.p, L->cnt.p, C,
                 L->U.p, nullptr, s);
  if (e != hipSuccess) return e;
  w.ham_exact(L->idx.p, L->cnt.p, L->U.p, L->ncell.p, L->Hy.p);
  hipLaunchKernelGGL(lw_accept_kernel, dim3(gb), dim3(256), 0, s, C, fin, L->Hy.p, L->cem.p,
                     L->accept.p, L->reason.p, L->cH.p, da ? L->cHs.p : nullptr,
                     da ? L->Hsy.p : nullptr);
  stat_add(w, &L->dstats.p->exact_e, L->cnt.p, 0, 1);
  if (w.use_sur) record_exact(w, fin, L->ncell.p);
  count_step(w, m, fin, 0, 1, 0, da ? 1 : 0);
  hipLaunchKernelGGL(lw_volume_commit_kernel, dim3(blocks_for((size_t)C * Nm * 3)), dim3(256),
                     0, s, w.ch, L->cpos.p, L->ccell.p, L->npos.p, L->ncell.p, L->accept.p);
  return hipGetLastError();
}

// HMC: leapfrog with the judge's forces (plain) or the surrogate's
// (screened), then the reference's test on the judge's energy, after the
// surrogate screen when screened.  Forces at the end of a leapfrog step are
// reused at the start of the next (the counters report the reference's
// 2 n_leapfrog + 1 evaluations, the statistics the evaluations made).
hipError_t hmc_step(WalkCtx& w, int t, int m, const jr_move& mv, bool da) {
  LearnedState* L = w.L;
  jr_handle* h = w.h;
  hipStream_t s = w.s;
  const int C = w.C, Nm = w.Nm;
  const size_t xs = (size_t)w.n_steps * 3 * Nm;
  const unsigned gw = blocks_for(C, 4), gc = blocks_for((size_t)C * Nm * 3);
  hipLaunchKernelGGL(lw_hmc_begin_kernel, dim3(gw), dim3(256), 0, s, w.ch,
                     w.normals + (size_t)t * 3 * Nm, xs, L->cpos.p, L->x0.p, L->mom.p,
                     L->K0.p);
  const int nl = mv.n_leapfrog;
  hipError_t e;
  auto eval = [&]() -> hipError_t {
    if (da)
      return run_sur(L, C, L->cpos.p, L->ccell.p, nullptr, nullptr, L->Us.p, L->F.p, s);
    return judge_eval(h, L, L->cpos.p, L->ccell.p, L->ctypes.p, L->cn.p, nullptr, nullptr, C,
                      L->U.p, L->F.p, s);
  };
  if ((e = eval()) != hipSuccess) return e;
  for (int l = 0; l < nl; ++l) {
    hipLaunchKernelGGL(lw_hmc_kick_drift_kernel, dim3(gc), dim3(256), 0, s, w.ch, L->mom.p,
                       L->cpos.p, L->F.p, w.step_sizes, m, l == 0 ? 1 : 2, 1);
    if ((e = eval()) != hipSuccess) return e;
  }
  hipLaunchKernelGGL(lw_hmc_kick_drift_kernel, dim3(gc), dim3(256), 0, s, w.ch, L->mom.p,
                     L->cpos.p, L->F.p, w.step_sizes, m, 1, 0);
  const unsigned gb = blocks_for(C);
  if (da) {
    w.ham_sur(nullptr, nullptr, L->Us.p, L->ccell.p, L->Hsy.p);
    hipLaunchKernelGGL(lw_stage1_kernel, dim3(gb), dim3(256), 0, s, C, nullptr, L->cHs.p,
                       L->Hsy.p, L->thr_screen.p, L->code.p);
    hipLaunchKernelGGL(lw_compact_kernel, dim3(1), dim3(1024), 0, s, C, L->code.p, 1, 0,
                       nullptr, nullptr, 0.0, L->idx.p, L->cnt.p);
    e = judge_eval(h, L, L->cpos.p, L->ccell.p, L->ctypes.p, L->cn.p, L->idx.p, L->cnt.p, C,
                   L->U.p, nullptr, s);
    if (e != hipSuccess) return e;
    w.ham_exact(L->idx.p, L->cnt.p, L->U.p, L->ccell.p, L->Hy.p);
    hipLaunchKernelGGL(lw_hmc_finish_kernel, dim3(gw), dim3(256), 0, s, w.ch, L->code.p,
                       L->Hy.p, L->cH.p, L->K0.p, L->mom.p, L->cem.p, L->accept.p, L->reason.p,
                       L->cpos.p, L->x0.p, L->cHs.p, L->Hsy.p, w.uniforms + (size_t)t * 4,
                       (size_t)w.n_steps * 4);
    record_exact(w, L->code.p, L->ccell.p);
    count_step(w, m, L->code.p, 0, 1, 0, 1);
    stat_add(w, &L->dstats.p->exact_e, L->cnt.p, 0, 1);
    stat_add(w, &L->dstats.p->sur_e, nullptr, C, nl + 1);
  } else {
    w.ham_exact(nullptr, nullptr, L->U.p, L->ccell.p, L->Hy.p);
    if (w.use_sur && w.record) {
      // the surrogate's prediction at the judged end point: held-out error
      e = run_sur(L, C, L->cpos.p, L->ccell.p, nullptr, nullptr, L->Us.p, nullptr, s);
      if (e != hipSuccess) return e;
      record_exact(w, nullptr, L->ccell.p);
      stat_add(w, &L->dstats.p->sur_e, nullptr, C, 1);
    }
    hipLaunchKernelGGL(lw_hmc_finish_kernel, dim3(gw), dim3(256), 0, s, w.ch, nullptr, L->Hy.p,
                       L->cH.p, L->K0.p, L->mom.p, L->cem.p, L->accept.p, L->reason.p,
                       L->cpos.p, L->x0.p, nullptr, nullptr, nullptr, (size_t)0);
    count_step(w, m, nullptr, 2 * nl + 1, 0, 2 * nl, 0);
    stat_add(w, &L->dstats.p->exact_fe, nullptr, C, nl + 1);
  }
  return hipGetLastError();
}

// Adam steps on minibatches drawn from the ring (all on the stream).
hipError_t train(jr_handle* h, LearnedState* L, hipStream_t s) {
  const int Bt = L->cfg.train_batch, Nm = L->Nmax;
  const int rows = Bt * Nm, chunks = (rows + 63) / 64;
  SurRowsArgs ra;
  ra.net = L->sur.dev();
  ra.n_rows = rows;
  ra.x = L->tX.p;
  ra.e = L->te.p;
  ra.c = L->tc.p;
  ra.gpart = L->gpart.p;
  for (int k = 0; k < L->cfg.train_steps; ++k) {
    hipLaunchKernelGGL(lw_train_draw_kernel, dim3(blocks_for(Bt)), dim3(256), 0, s, Bt, L->rs.p,
                       (unsigned long long)h->cfg.seed, L->train_counter++, L->tidx.p);
    hipLaunchKernelGGL(lw_train_gather_kernel, dim3(blocks_for((size_t)rows * SG_D)), dim3(256),
                       0, s, Bt, Nm, L->tidx.p, L->ring.p, L->ringE.p, L->ringN.p, L->tX.p,
                       L->tE.p, L->tn.p);
    hipLaunchKernelGGL(sur_rows_forward_kernel, dim3(std::min(chunks, 4 * L->n_cu)), dim3(64), 0,
                       s, ra);
    hipLaunchKernelGGL(lw_loss_weights_kernel, dim3(Bt), dim3(256), 0, s, Bt, Nm, L->te.p,
                       L->tE.p, L->tn.p, L->tc.p, L->tsq.p);
    hipLaunchKernelGGL(sur_rows_backward_kernel, dim3(chunks), dim3(64), 0, s, ra);
    hipLaunchKernelGGL(sur_reduce_kernel, dim3((SG_NP + 255) / 256), dim3(256), 0, s, L->gpart.p,
                       chunks, L->grad.p);
    L->adam_t += 1;
    const float c1 = (float)(1.0 - std::pow(0.9, (double)L->adam_t));
    const float c2 = (float)(1.0 - std::pow(0.999, (double)L->adam_t));
    hipLaunchKernelGGL(sur_adam_kernel, dim3((SG_NP + 255) / 256), dim3(256), 0, s, L->sur.P.p,
                       L->adam_m.p, L->adam_v.p, L->grad.p, (float)L->cfg.lr, c1, c2);
    L->sur.cast(s);
  }
  if (L->cfg.train_steps > 0)
    hipLaunchKernelGGL(lw_loss_sum_kernel, dim3(1), dim3(64), 0, s, Bt, L->tsq.p, L->dstats.p);
  return hipGetLastError();
}

}  // namespace

// Error model of the surrogate from one iteration's held-out sums (per-atom
// units e = (U_s - U)/N, v = V/N): e = ca + cv v fitted over all replicas by
// least squares, then per replica the residuals' mean (bias) and spread.
// Needs min_heldout evaluations; otherwise the previous model stays.
static void fit_error_model(LearnedState* L, const std::vector<double>& hs) {
  const int R = (int)L->bias.size();
  double t[LW_HS] = {0, 0, 0, 0, 0, 0};
  for (int r = 0; r < R; ++r)
    for (int q = 0; q < LW_HS; ++q) t[q] += hs[(size_t)r * LW_HS + q];
  const double n = t[0];
  if (n < std::max(1, L->cfg.min_heldout)) return;
  const double mv = t[3] / n, me = t[1] / n;
  const double var_v = t[4] / n - mv * mv, cov = t[5] / n - me * mv;
  const double cv = var_v > 1e-12 * std::max(1.0, mv * mv) ? cov / var_v : 0.0;
  const double ca = me - cv * mv;
  auto resid = [&](const double* h, double& s1, double& s2) {
    s1 = h[1] - ca * h[0] - cv * h[3];
    s2 = h[2] + ca * ca * h[0] + cv * cv * h[4] - 2.0 * ca * h[1] - 2.0 * cv * h[5] +
         2.0 * ca * cv * h[3];
  };
  double g1, g2;
  resid(t, g1, g2);
  const double gb = g1 / n, gs = std::sqrt(std::max(g2 / n - gb * gb, 0.0));
  for (int r = 0; r < R; ++r) {
    const double* h = &hs[(size_t)r * LW_HS];
    if (h[0] >= 2.0) {
      double s1, s2;
      resid(h, s1, s2);
      L->bias[r] = s1 / h[0];
      L->sig[r] = std::sqrt(std::max(s2 / h[0] - L->bias[r] * L->bias[r], 0.0));
    } else {
      L->bias[r] = gb;
      L->sig[r] = gs;
    }
  }
  L->ca = ca;
  L->cv = cv;
  L->rmse = std::sqrt(std::max(g2, 0.0) / n);
  L->rmse_n = (long long)n;
}

int learned_walk(jr_handle* h) {
  LearnedState* L = h->learned;
  const auto& c = h->cfg;
  hipStream_t s = h->stream;
  WalkCtx w;
  w.h = h;
  w.L = L;
  w.s = s;
  w.C = L->C;
  w.Nm = L->Nmax;
  w.n_steps = c.n_mcmc_steps;
  w.use_sur = L->mode != JR_LEARNED_PLAIN;
  w.cap = L->cfg.train_max_epa;
  w.ch = LWChains{L->C, h->n_walk, L->Nmax, c.n_moves, L->cn.p};
  w.normals = h->normals.p;
  w.uniforms = h->uniforms.p;
  w.step_sizes = h->step_sizes.p;
  w.counters = h->counters.p;
  for (double mu : h->h_mu)
    if (mu != 0.0) return h->fail(JR_E_UNSUPPORTED, "the learned walk is NPT (mu = 0)");
  HIPCHK(h, hipEventRecord(L->ev[0], s));
  LWGatherArgs ga;
  ga.C = L->C;
  ga.n_walk = h->n_walk;
  ga.K = c.n_walkers;
  ga.Nmax = L->Nmax;
  ga.walk = h->walk.p;
  ga.pop = PopPtrs{h->pos.p, h->types.p, h->n_atoms.p, h->cell.p, h->energy.p, h->dir.p};
  ga.emax = h->emax.p;
  ga.pressure = h->pressure.p;
  ga.cpos = L->cpos.p;
  ga.cdir = L->cdir.p;
  ga.ccell = L->ccell.p;
  ga.cH = L->cH.p;
  ga.cem = L->cem.p;
  ga.cpres = L->cpres.p;
  ga.ctypes = L->ctypes.p;
  ga.cn = L->cn.p;
  hipLaunchKernelGGL(lw_gather_kernel, dim3(L->C), dim3(256), 0, s, ga);
  hipLaunchKernelGGL(lw_contour_kernel, dim3(1), dim3(256), 0, s, L->C, L->cem.p, L->cpres.p,
                     L->ccell.p, L->cn.p, L->vol_lo, L->vol_hi, L->dcontour.p);
  // the shared move schedule decides which launches follow
  double contour_epa = INFINITY;
  HIPCHK(h, hipMemcpyAsync(L->sched.data(), h->sched.p, sizeof(int) * c.n_mcmc_steps,
                           hipMemcpyDeviceToHost, s));
  HIPCHK(h, hipMemcpyAsync(&contour_epa, L->dcontour.p, sizeof(double), hipMemcpyDeviceToHost,
                           s));
  HIPCHK(h, hipStreamSynchronize(s));
  // auto: the held-out error meets the bound and no state the contours admit
  // lies above the energy range the surrogate trains on
  const bool da = L->mode == JR_LEARNED_SCREENED ||
                  (L->mode == JR_LEARNED_AUTO && L->rmse <= L->cfg.rmse_bound &&
                   L->rmse_n >= L->cfg.min_heldout && contour_epa <= L->cfg.train_max_epa);
  if (da) {
    // contours of this iteration from the error model of the last one
    HIPCHK(h, hipMemcpyAsync(L->dbias.p, L->bias.data(), sizeof(double) * L->bias.size(),
                             hipMemcpyHostToDevice, s));
    HIPCHK(h, hipMemcpyAsync(L->dsig.p, L->sig.data(), sizeof(double) * L->sig.size(),
                             hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(lw_thr_kernel, dim3(blocks_for(L->C)), dim3(256), 0, s, L->C, h->n_walk,
                       L->cem.p, L->cn.p, L->dbias.p, L->dsig.p, L->cfg.delta_k,
                       L->thr_steer.p, L->thr_screen.p);
    // the surrogate's verdict on every chain's current state, with the
    // parameters this iteration uses throughout
    HIPCHK(h, run_sur(L, L->C, L->cpos.p, L->ccell.p, nullptr, nullptr, L->Us.p, nullptr, s));
    w.ham_sur(nullptr, nullptr, L->Us.p, L->ccell.p, L->cHs.p);
    hipLaunchKernelGGL(lw_stat_add_kernel, dim3(1), dim3(64), 0, s, &L->dstats.p->sur_e,
                       nullptr, (long long)L->C, 1LL);
  }
  for (int t = 0; t < c.n_mcmc_steps; ++t) {
    const int m = L->sched[t];
    const jr_move& mv = L->moves[m];
    hipError_t e = mv.type == JR_MOVE_GMC   ? gmc_step(w, t, m, mv, da)
                   : mv.type == JR_MOVE_HMC ? hmc_step(w, t, m, mv, da)
                                            : volume_step(w, t, m, mv, da);
    if (e != hipSuccess)
      return h->fail(JR_E_HIP, "learned walk step %d: %s", t, hipGetErrorString(e));
  }
  hipLaunchKernelGGL(lw_scatter_kernel, dim3(L->C), dim3(256), 0, s, ga);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipEventRecord(L->ev[1], s));
  const bool do_train = w.use_sur && L->stored >= L->cfg.min_train && L->cfg.train_steps > 0;
  HIPCHK(h, hipEventRecord(L->ev[2], s));
  if (do_train) HIPCHK(h, train(h, L, s));
  HIPCHK(h, hipEventRecord(L->ev[3], s));
  // statistics of this iteration (one sync)
  LWStats st;
  RingState rs;
  int of = 0;
  std::vector<double> hs(L->hs.n);
  HIPCHK(h, hipMemcpyAsync(hs.data(), L->hs.p, sizeof(double) * hs.size(), hipMemcpyDeviceToHost,
                           s));
  HIPCHK(h, hipMemcpyAsync(&st, L->dstats.p, sizeof(st), hipMemcpyDeviceToHost, s));
  HIPCHK(h, hipMemcpyAsync(&rs, L->rs.p, sizeof(rs), hipMemcpyDeviceToHost, s));
  HIPCHK(h, hipMemcpyAsync(&of, L->overflow.p, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(h, zero_stats(L, s));
  HIPCHK(h, hipStreamSynchronize(s));
  if (of) return h->fail(JR_E_UNSUPPORTED, "judge neighbour list overflow (> %d)", MP_MAXNB);
  float ms_walk = 0.f, ms_train = 0.f;
  HIPCHK(h, hipEventElapsedTime(&ms_walk, L->ev[0], L->ev[1]));
  HIPCHK(h, hipEventElapsedTime(&ms_train, L->ev[2], L->ev[3]));
  jr_learned_stats& a = L->acc;
  (da ? a.iterations_screened : a.iterations_plain) += 1;
  a.proposals += (int64_t)st.proposals;
  a.stage1_pass += (int64_t)st.stage1_pass;
  a.accepted += (int64_t)st.accepted;
  a.exact_force_evals += (int64_t)st.exact_fe;
  a.exact_energy_evals += (int64_t)st.exact_e;
  a.surrogate_evals += (int64_t)st.sur_e;
  if (do_train) {
    a.train_steps += L->cfg.train_steps;
    a.train_loss = st.loss;
  }
  a.walk_ms += ms_walk;
  a.train_ms += ms_train;
  L->stored = rs.size;
  a.stored = rs.size;
  fit_error_model(L, hs);
  a.heldout_rmse = L->rmse;
  return JR_OK;
}

namespace {

// Gathers the walkers walk [R][nw] (indices within their replica) into the
// first R nw chains under the contours emax [R], then sets up the walk's
// screening for them: whether it screens (the walk's rule), and if so the
// thresholds and the surrogate's verdict on their current states.  One sync.
int gather_chains(jr_handle* h, int nw, const int* walk, const double* emax, bool* da_out) {
  LearnedState* L = h->learned;
  const auto& c = h->cfg;
  hipStream_t s = h->stream;
  const int Ct = c.n_replicas * nw;
  LWGatherArgs ga;
  ga.C = Ct;
  ga.n_walk = nw;
  ga.K = c.n_walkers;
  ga.Nmax = L->Nmax;
  ga.walk = walk;
  ga.pop = PopPtrs{h->pos.p, h->types.p, h->n_atoms.p, h->cell.p, h->energy.p, h->dir.p};
  ga.emax = emax;
  ga.pressure = h->pressure.p;
  ga.cpos = L->cpos.p;
  ga.cdir = L->cdir.p;
  ga.ccell = L->ccell.p;
  ga.cH = L->cH.p;
  ga.cem = L->cem.p;
  ga.cpres = L->cpres.p;
  ga.ctypes = L->ctypes.p;
  ga.cn = L->cn.p;
  hipLaunchKernelGGL(lw_gather_kernel, dim3(Ct), dim3(256), 0, s, ga);
  hipLaunchKernelGGL(lw_contour_kernel, dim3(1), dim3(256), 0, s, Ct, L->cem.p, L->cpres.p,
                     L->ccell.p, L->cn.p, L->vol_lo, L->vol_hi, L->dcontour.p);
  double contour_epa = INFINITY;
  HIPCHK(h, hipMemcpyAsync(&contour_epa, L->dcontour.p, sizeof(double), hipMemcpyDeviceToHost,
                           s));
  HIPCHK(h, hipStreamSynchronize(s));
  const bool da = L->mode == JR_LEARNED_SCREENED ||
                  (L->mode == JR_LEARNED_AUTO && L->rmse <= L->cfg.rmse_bound &&
                   L->rmse_n >= L->cfg.min_heldout && contour_epa <= L->cfg.train_max_epa);
  if (da) {
    HIPCHK(h, hipMemcpyAsync(L->dbias.p, L->bias.data(), sizeof(double) * L->bias.size(),
                             hipMemcpyHostToDevice, s));
    HIPCHK(h, hipMemcpyAsync(L->dsig.p, L->sig.data(), sizeof(double) * L->sig.size(),
                             hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(lw_thr_kernel, dim3(blocks_for(Ct)), dim3(256), 0, s, Ct, nw, L->cem.p,
                       L->cn.p, L->dbias.p, L->dsig.p, L->cfg.delta_k, L->thr_steer.p,
                       L->thr_screen.p);
    HIPCHK(h, run_sur(L, Ct, L->cpos.p, L->ccell.p, nullptr, nullptr, L->Us.p, nullptr, s));
    hipLaunchKernelGGL(lw_ham_kernel, dim3(blocks_for(Ct)), dim3(256), 0, s, Ct, nullptr,
                       nullptr, L->Us.p, L->ccell.p, L->cpres.p, L->cn.p, L->ca, L->cv,
                       L->cHs.p);
    HIPCHK(h, hipGetLastError());
  }
  *da_out = da;
  return JR_OK;
}

// A walk context over the first R nw chains with the given draws, step sizes
// and counters, without statistics or training records.
WalkCtx side_ctx(jr_handle* h, int nw, int n_steps, const double* normals,
                 const double* uniforms, const double* step_sizes, int* counters) {
  LearnedState* L = h->learned;
  WalkCtx w;
  w.h = h;
  w.L = L;
  w.s = h->stream;
  w.C = h->cfg.n_replicas * nw;
  w.Nm = L->Nmax;
  w.n_steps = n_steps;
  w.use_sur = L->mode != JR_LEARNED_PLAIN;
  w.cap = L->cfg.train_max_epa;
  w.ch = LWChains{w.C, nw, L->Nmax, h->cfg.n_moves, L->cn.p};
  w.normals = normals;
  w.uniforms = uniforms;
  w.step_sizes = step_sizes;
  w.counters = counters;
  w.stats = w.record = false;
  return w;
}

hipError_t run_step(WalkCtx& w, int t, int m, bool da) {
  const jr_move& mv = w.L->moves[m];
  return mv.type == JR_MOVE_GMC   ? gmc_step(w, t, m, mv, da)
         : mv.type == JR_MOVE_HMC ? hmc_step(w, t, m, mv, da)
                                  : volume_step(w, t, m, mv, da);
}

// The judge's neighbour-list overflow flag (read and cleared; one sync).
int check_overflow(jr_handle* h) {
  LearnedState* L = h->learned;
  int of = 0;
  HIPCHK(h, hipMemcpyAsync(&of, L->overflow.p, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (!of) return JR_OK;
  HIPCHK(h, hipMemsetAsync(L->overflow.p, 0, sizeof(int), h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return h->fail(JR_E_UNSUPPORTED, "judge neighbour list overflow (> %d)", MP_MAXNB);
}

int npt_only(jr_handle* h) {
  for (double mu : h->h_mu)
    if (mu != 0.0) return h->fail(JR_E_UNSUPPORTED, "the learned walk is NPT (mu = 0)");
  return JR_OK;
}

}  // namespace

int learned_chains(jr_handle* h) { return h->learned->C; }

// Trial moves of the learned walk (sampling/mwg.py per_move_fns as the
// step-size adaptation uses them): move m once from each sampled walker
// (sample_idx [R][n] on the host) at trial step size tstep [R] under the
// contours temax [R], with the draws tnorm [R][n][3][Nmax] and tunif
// [R][n][4] (device), in the walk's mode (the surrogate steers and screens
// when the walk would); reasons [R][n] to the host (0 = accepted).  The
// population, the counters, the statistics and the training ring are left
// as they are.  Passes of at most the walk's chain count C.
int learned_trial_moves(jr_handle* h, int m, int n, const int32_t* sample_idx,
                        const double* tstep, const double* temax, const double* tnorm,
                        const double* tunif, int32_t* reason) {
  LearnedState* L = h->learned;
  const auto& c = h->cfg;
  hipStream_t s = h->stream;
  const int R = c.n_replicas, Nm = L->Nmax, nm = c.n_moves;
  if (int rc = npt_only(h)) return rc;
  const size_t C = (size_t)L->C;
  if (!L->trial_idx.p) {
    HIPCHK(h, L->trial_idx.alloc(C));
    HIPCHK(h, L->trial_norm.alloc(C * 3 * Nm));
    HIPCHK(h, L->trial_unif.alloc(C * 4));
    HIPCHK(h, L->trial_steps.alloc((size_t)R * nm));
  }
  // the step sizes with the trial move's column replaced
  HIPCHK(h, hipMemcpyAsync(L->trial_steps.p, h->step_sizes.p, sizeof(double) * R * nm,
                           hipMemcpyDeviceToDevice, s));
  HIPCHK(h, hipMemcpy2DAsync(L->trial_steps.p + m, sizeof(double) * nm, tstep, sizeof(double),
                             sizeof(double), R, hipMemcpyDeviceToDevice, s));
  const int per = (int)(C / R);  // sampled walkers per replica in one pass
  std::vector<int> rs(C);
  L->trial_host.resize(C);
  for (int j0 = 0; j0 < n; j0 += per) {
    const int nc = std::min(per, n - j0), Ct = R * nc;
    for (int r = 0; r < R; ++r)
      for (int j = 0; j < nc; ++j)
        L->trial_host[(size_t)r * nc + j] = sample_idx[(size_t)r * n + j0 + j];
    HIPCHK(h, hipMemcpyAsync(L->trial_idx.p, L->trial_host.data(), sizeof(int) * Ct,
                             hipMemcpyHostToDevice, s));
    const size_t row = (size_t)3 * Nm;
    HIPCHK(h, hipMemcpy2DAsync(L->trial_norm.p, sizeof(double) * row * nc, tnorm + j0 * row,
                               sizeof(double) * row * n, sizeof(double) * row * nc, R,
                               hipMemcpyDeviceToDevice, s));
    HIPCHK(h, hipMemcpy2DAsync(L->trial_unif.p, sizeof(double) * 4 * nc,
                               tunif + (size_t)j0 * 4, sizeof(double) * 4 * n,
                               sizeof(double) * 4 * nc, R, hipMemcpyDeviceToDevice, s));
    bool da = false;
    if (int rc = gather_chains(h, nc, L->trial_idx.p, temax, &da)) return rc;
    WalkCtx w = side_ctx(h, nc, 1, L->trial_norm.p, L->trial_unif.p, L->trial_steps.p,
                         nullptr);
    hipError_t e = run_step(w, 0, m, da);
    if (e != hipSuccess)
      return h->fail(JR_E_HIP, "learned trial move: %s", hipGetErrorString(e));
    HIPCHK(h, hipMemcpyAsync(rs.data(), L->reason.p, sizeof(int) * Ct, hipMemcpyDeviceToHost,
                             s));
    if (int rc = check_overflow(h)) return rc;
    for (int r = 0; r < R; ++r)
      for (int j = 0; j < nc; ++j)
        reason[(size_t)r * n + j0 + j] = rs[(size_t)r * nc + j];
  }
  return JR_OK;
}

// Burn-in walk of the learned handle (init/burn_in.py:48-148) for the
// walkers walk [R][nw] (device, nw <= C / R): wl sub-steps with the move
// schedule sched (host) at the fixed contours emax [R], the draws normals
// [R][nw][wl][3][Nmax] and uniforms [R][nw][wl][4] and the handle's step
// sizes, as the walk takes them (judge-steered, or surrogate-steered and
// screened when the walk screens); move counters added into ctr
// [R][n_moves][JR_NCOUNT]; the walkers are written back.  No statistics or
// training records.
int learned_burn(jr_handle* h, int wl, const int* sched, int nw, const int* walk,
                 const double* emax, const double* normals, const double* uniforms,
                 int* ctr) {
  LearnedState* L = h->learned;
  if (int rc = npt_only(h)) return rc;
  if (nw < 1 || (size_t)h->cfg.n_replicas * nw > (size_t)L->C)
    return h->fail(JR_E_ARG, "burn-in pass of %d walkers per replica exceeds the chains", nw);
  bool da = false;
  if (int rc = gather_chains(h, nw, walk, emax, &da)) return rc;
  WalkCtx w = side_ctx(h, nw, wl, normals, uniforms, h->step_sizes.p, ctr);
  for (int t = 0; t < wl; ++t) {
    hipError_t e = run_step(w, t, sched[t], da);
    if (e != hipSuccess)
      return h->fail(JR_E_HIP, "learned burn-in step %d: %s", t, hipGetErrorString(e));
  }
  LWGatherArgs ga;
  std::memset(&ga, 0, sizeof(ga));
  ga.C = w.C;
  ga.n_walk = nw;
  ga.K = h->cfg.n_walkers;
  ga.Nmax = L->Nmax;
  ga.walk = walk;
  ga.pop = PopPtrs{h->pos.p, h->types.p, h->n_atoms.p, h->cell.p, h->energy.p, h->dir.p};
  ga.cpos = L->cpos.p;
  ga.cdir = L->cdir.p;
  ga.ccell = L->ccell.p;
  ga.cH = L->cH.p;
  ga.cn = L->cn.p;
  hipLaunchKernelGGL(lw_scatter_kernel, dim3(w.C), dim3(256), 0, h->stream, ga);
  HIPCHK(h, hipGetLastError());
  return check_overflow(h);
}

}  // namespace jrl

using namespace jrl;

extern "C" int jr_learned_get_stats(jr_handle* h, jr_learned_stats* out, int32_t reset) {
  if (!h || !out) return set_global_err(JR_E_ARG, "NULL argument");
  if (!h->learned) return h->fail(JR_E_STATE, "the handle has no learned model");
  *out = h->learned->acc;
  out->heldout_rmse = h->learned->rmse;
  out->stored = h->learned->stored;
  if (reset) {
    const double r = h->learned->acc.heldout_rmse, l = h->learned->acc.train_loss;
    h->learned->acc = jr_learned_stats{};
    h->learned->acc.heldout_rmse = r;
    h->learned->acc.train_loss = l;
  }
  return JR_OK;
}

extern "C" int jr_learned_get_params(jr_handle* h, float* params) {
  if (!h || !params) return set_global_err(JR_E_ARG, "NULL argument");
  LearnedState* L = h->learned;
  if (!L || !L->sur.P.p) return h->fail(JR_E_STATE, "the handle has no surrogate");
  HIPCHK(h, hipSetDevice(h->cfg.device));
  HIPCHK(h, hipMemcpyAsync(params, L->sur.P.p, sizeof(float) * SG_NP, hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return JR_OK;
}

extern "C" int jr_learned_set_params(jr_handle* h, const float* params, int32_t reset_adam) {
  if (!h || !params) return set_global_err(JR_E_ARG, "NULL argument");
  LearnedState* L = h->learned;
  if (!L || !L->sur.P.p) return h->fail(JR_E_STATE, "the handle has no surrogate");
  HIPCHK(h, hipSetDevice(h->cfg.device));
  hipStream_t s = h->stream;
  HIPCHK(h, hipMemcpyAsync(L->sur.P.p, params, sizeof(float) * SG_NP, hipMemcpyHostToDevice, s));
  L->sur.cast(s);
  HIPCHK(h, hipGetLastError());
  if (reset_adam) {
    HIPCHK(h, hipMemsetAsync(L->adam_m.p, 0, sizeof(float) * SG_NP, s));
    HIPCHK(h, hipMemsetAsync(L->adam_v.p, 0, sizeof(float) * SG_NP, s));
    L->adam_t = 0;
  }
  // a new surrogate has no held-out record yet
  L->rmse = INFINITY;
  L->rmse_n = 0;
  L->ca = L->cv = 0.0;
  std::fill(L->bias.begin(), L->bias.end(), 0.0);
  std::fill(L->sig.begin(), L->sig.end(), 0.0);
  HIPCHK(h, hipStreamSynchronize(s));
  return JR_OK;
}

extern "C" int jr_learned_set_mode(jr_handle* h, int32_t mode) {
  if (!h) return set_global_err(JR_E_ARG, "NULL handle");
  LearnedState* L = h->learned;
  if (!L) return h->fail(JR_E_STATE, "the handle has no learned model");
  if (mode < JR_LEARNED_PLAIN || mode > JR_LEARNED_SCREENED)
    return h->fail(JR_E_ARG, "mode %d unknown", mode);
  if (mode != JR_LEARNED_PLAIN && !L->sur.P.p)
    return h->fail(JR_E_STATE, "the handle was created without a surrogate");
  L->mode = mode;
  return JR_OK;
}
