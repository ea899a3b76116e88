// Host side of the learned part of libjrhip: the message-passing judge
// (batched energy and forces).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "jr_host.h"
#include "jr_mpnn.h"

using namespace jr;
using namespace jrh;

namespace {

constexpr size_t kMpLds =
    sizeof(float) * (4 * MP_THREADS + 2 * MP_THREADS * MP_C +
                     MP_THREADS * MP_NB * MP_MAXS + MP_NB * MP_C);

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

// Launch the judge on items [0, a.B) with `grid` persistent blocks.
hipError_t launch_mpnn(const MPNNArgs& a, const float* w, int grid, hipStream_t s) {
  hipError_t e;
  if (a.net.S == 1) {
    e = hipFuncSetAttribute((const void*)mpnn_kernel<1>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMpLds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(mpnn_kernel<1>, dim3(grid), dim3(MP_THREADS), kMpLds, s, a, w);
  } else {
    e = hipFuncSetAttribute((const void*)mpnn_kernel<2>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMpLds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(mpnn_kernel<2>, dim3(grid), dim3(MP_THREADS), kMpLds, s, a, w);
  }
  return hipGetLastError();
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
  const int grid = std::min(B, device_cus(device));
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

constexpr size_t kSgLds = sizeof(float) * 4 * 256 + sizeof(__bf16) * 256 * (SG_D + SG_H);

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

hipError_t launch_sur_eval(const SurEvalArgs& a, int grid, hipStream_t s) {
  hipError_t e = hipFuncSetAttribute((const void*)sur_eval_kernel,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)kSgLds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(sur_eval_kernel, dim3(grid), dim3(256), kSgLds, s, a);
  return hipGetLastError();
}

struct StreamHolder {
  hipStream_t s = nullptr;
  ~StreamHolder() {
    if (s) (void)hipStreamDestroy(s);
  }
};

}  // namespace

extern "C" int jr_sur_eval(const float* params, int32_t device, int32_t B, int32_t Nmax,
                           const double* pos, const int32_t* types,
                           const int32_t* n_atoms, const double* cell, int32_t n_species,
                           double* U, double* F, uint16_t* x_out) {
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
  HIPCHK_G(hipSetDevice(device));
  StreamHolder sh;
  HIPCHK_G(hipStreamCreate(&sh.s));
  hipStream_t s = sh.s;
  SurParams w;
  HIPCHK_G(w.alloc());
  HIPCHK_G(hipMemcpyAsync(w.P.p, params, sizeof(float) * SG_NP, hipMemcpyHostToDevice, s));
  w.cast(s);
  HIPCHK_G(hipGetLastError());
  DBuf<double> d_pos, d_cell, d_U, d_F;
  DBuf<int> d_t, d_n;
  DBuf<__bf16> d_x;
  HIPCHK_G(d_pos.alloc((size_t)B * Nmax * 3));
  HIPCHK_G(d_cell.alloc((size_t)B * 9));
  HIPCHK_G(d_t.alloc((size_t)B * Nmax));
  HIPCHK_G(d_n.alloc((size_t)B));
  HIPCHK_G(d_U.alloc((size_t)B));
  if (F) HIPCHK_G(d_F.alloc((size_t)B * Nmax * 3));
  if (x_out) HIPCHK_G(d_x.alloc((size_t)B * Nmax * SG_D));
  HIPCHK_G(hipMemcpyAsync(d_pos.p, pos, sizeof(double) * B * Nmax * 3, hipMemcpyHostToDevice, s));
  HIPCHK_G(hipMemcpyAsync(d_cell.p, cell, sizeof(double) * B * 9, hipMemcpyHostToDevice, s));
  HIPCHK_G(hipMemcpyAsync(d_t.p, types, sizeof(int) * B * Nmax, hipMemcpyHostToDevice, s));
  HIPCHK_G(hipMemcpyAsync(d_n.p, n_atoms, sizeof(int) * B, hipMemcpyHostToDevice, s));
  if (F) HIPCHK_G(hipMemsetAsync(d_F.p, 0, sizeof(double) * B * Nmax * 3, s));
  if (x_out) HIPCHK_G(hipMemsetAsync(d_x.p, 0, sizeof(__bf16) * B * Nmax * SG_D, s));
  SurEvalArgs a;
  a.net = w.dev();
  a.B = B;
  a.Nmax = Nmax;
  a.S = n_species;
  a.pos = d_pos.p;
  a.types = d_t.p;
  a.n_atoms = d_n.p;
  a.cell = d_cell.p;
  a.cfg_index = nullptr;
  a.E = d_U.p;
  a.F = F ? d_F.p : nullptr;
  a.xout = x_out ? d_x.p : nullptr;
  HIPCHK_G(launch_sur_eval(a, std::min(B, device_cus(device)), s));
  if (U) HIPCHK_G(hipMemcpyAsync(U, d_U.p, sizeof(double) * B, hipMemcpyDeviceToHost, s));
  if (F)
    HIPCHK_G(hipMemcpyAsync(F, d_F.p, sizeof(double) * B * Nmax * 3, hipMemcpyDeviceToHost, s));
  if (x_out)
    HIPCHK_G(hipMemcpyAsync(x_out, d_x.p, sizeof(__bf16) * B * Nmax * SG_D,
                            hipMemcpyDeviceToHost, s));
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
