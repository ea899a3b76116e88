// Message-passing judge (seeded random-init invariant MPNN), FP32: energy
// and analytic forces of one cubic-cell configuration per 256-thread
// workgroup, one thread per receiving atom.
//
//   b_n(r) = sqrt(2/rc) sin(n pi r/rc)/r * u(r/rc),  u(x) = 1-28x^6+48x^7-21x^8
//   S_i[n][s] = sum_{j in N(i), t_j = s} b_n(r_ij)
//   h1_i = emb[t_i] + silu(m1_i U1 + c1),  m1_i[c] = sum_n W1[n][c] sum_s emb[s][c] S_i[n][s]
//   h2_i = h1_i + silu(m2_i U2 + c2),      m2_i[c] = sum_n W2[n][c] sum_j b_n(r_ij) h1_j[c]
//   E = sum_i wo.h2_i + 1/2 sum_{i, j in N(i), r < core_rc} phi(r),
//   phi(r) = core_eps ((core_sig/r)^12 - (core_sig/core_rc)^12)
//
// N(i) holds every periodic image within rc; with L >= rc (checked on the
// host) those are, per axis, the minimum image and the next one on the other
// side, so a pair (i, j) has at most 8 images.  A per-thread list of
// (j, image) entries is built once per evaluation (k-major in a global
// scratch buffer, coalesced), then walked by the passes:
//   1. S_i and the core energy; h1_i into LDS
//   2. sum_j b_n h1_j in 16-channel chunks -> m2_i, E_i; M_i = dE/dm2_i into LDS
//   3. per pair g = dE/dr over both directed edges through layer 2
//      (sum_c W2[n][c] (M_i[c] h1_j[c] + M_j[c] h1_i[c]) b'_n) and
//      dE/dh1_i = wo + sum_n W2[n] (.) sum_j b_n M_j  (16-channel chunks)
//   4. layer 1 (dE/dS of both ends) and core terms; F_i = sum_pairs g d/r.
#pragma once

#include <hip/hip_runtime.h>

#include "jr_images.h"

// FP32 judge: multiply-adds contract to FMA here (the library's FP64 walk
// keeps -ffp-contract=off for its canonical order; this file is compared
// with the FP64 oracle to a stated tolerance, not bit for bit).
#pragma clang fp contract(fast)

namespace jr {

constexpr int MP_C = 64;        // channels
constexpr int MP_NB = 8;        // radial functions
constexpr int MP_MAXS = 2;      // species (LDS budget: 148 KB per block)
constexpr int MP_THREADS = 256; // atoms per configuration (max)
constexpr int MP_MAXNB = 1024;  // neighbour-list capacity per atom
constexpr int MP_CH = 16;       // channel chunk

struct MPNNDev {
  int S;
  float rc, core_eps, core_sig, core_rc;
  float pref;     // sqrt(2 / rc)
  float core_sh;  // (core_sig / core_rc)^12
  float inv_rc;   // 1 / rc
  float pi_rc;    // pi / rc
};

// Packed weight buffer: emb [S][C], W1 [NB][C], U1 [C][C], c1 [C],
// W2 [NB][C], U2 [C][C], c2 [C], wo [C].
template <int S>
struct MPW {
  static constexpr int emb = 0;
  static constexpr int W1 = emb + S * MP_C;
  static constexpr int U1 = W1 + MP_NB * MP_C;
  static constexpr int c1 = U1 + MP_C * MP_C;
  static constexpr int W2 = c1 + MP_C;
  static constexpr int U2 = W2 + MP_NB * MP_C;
  static constexpr int c2 = U2 + MP_C * MP_C;
  static constexpr int wo = c2 + MP_C;
  static constexpr int total = wo + MP_C;
};

struct MPNNArgs {
  MPNNDev net;
  int B, Nmax;
  const double* pos;     // [B][Nmax][3]
  const int* types;      // [B][Nmax]
  const int* n_atoms;    // [B]
  const double* cell;    // [B][9] cubic
  const int* cfg_index;  // [B] configuration slot per item, or NULL (= item)
  const int* count;      // device item count (<= B), or NULL (= B)
  double* E;             // [B] model energy
  double* F;             // [B][Nmax][3] or NULL
  unsigned* nbl;         // scratch [gridDim.x][MP_MAXNB][MP_THREADS]
  int* overflow;         // set to 1 if a list exceeds MP_MAXNB
};

__device__ __forceinline__ float silu_f(float z) {
  return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}
__device__ __forceinline__ float silu_d_f(float z) {
  const float s = __builtin_amdgcn_rcpf(1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// Basis b_n (and b'_n when D) at distance r (ir = 1/r).  sin/cos of
// pi r/rc come from the hardware v_sin/v_cos (argument in revolutions), the
// higher orders from the angle-addition recurrence.
template <bool D>
__device__ __forceinline__ void mp_basis(float r, float ir, const MPNNDev& P,
                                         float (&b)[MP_NB], float (&db)[MP_NB]) {
  const float x = r * P.inv_rc;
  const float x2 = x * x, x5 = x2 * x2 * x, x6 = x5 * x, x7 = x6 * x;
  const float u = 1.0f - 28.0f * x6 + 48.0f * x7 - 21.0f * x7 * x;
  const float du = (-168.0f * x5 + 336.0f * x6 - 168.0f * x7) * P.inv_rc;
  const float s1 = __builtin_amdgcn_sinf(0.5f * x);
  const float c1 = __builtin_amdgcn_cosf(0.5f * x);
  const float pu = P.pref * ir * u;
  float s = s1, c = c1;
#pragma unroll
  for (int n = 0; n < MP_NB; ++n) {
    b[n] = pu * s;
    if (D) {
      const float a = (float)(n + 1) * P.pi_rc;
      db[n] = P.pref * ir * (a * c * u + s * (du - u * ir));
    }
    const float sn = s * c1 + c * s1, cn = c * c1 - s * s1;
    s = sn;
    c = cn;
  }
}

// Image displacement d (fractional units) of list entry e from atom i.
__device__ __forceinline__ void mp_edge(unsigned e, const float4* U, float4 ui,
                                        float& dx, float& dy, float& dz,
                                        int& j, int& tj) {
  j = (int)(e & 511u);
  const float4 uj = U[j];
  dx = uj.x - ui.x;
  dy = uj.y - ui.y;
  dz = uj.z - ui.z;
  dx -= rintf(dx);
  dy -= rintf(dy);
  dz -= rintf(dz);
  if (e & 512u) dx -= copysignf(1.0f, dx);
  if (e & 1024u) dy -= copysignf(1.0f, dy);
  if (e & 2048u) dz -= copysignf(1.0f, dz);
  tj = __float_as_int(uj.w);
}

// LDS of the judge: positions and h1 for the energy; with forces also M,
// the layer-1 gradients and W2 (bytes).
constexpr size_t mp_lds_bytes(bool forces) {
  return sizeof(float) * (4 * MP_THREADS + MP_THREADS * MP_C +
                          (forces ? MP_THREADS * MP_C + MP_THREADS * MP_NB * MP_MAXS + MP_NB * MP_C
                                  : 0));
}

// FORCES = false is a separate instantiation: without the backward its
// registers and LDS (68 KB) allow two blocks per CU.
template <int S, bool FORCES>
__global__ void __launch_bounds__(MP_THREADS)
    mpnn_kernel(MPNNArgs a, const float* __restrict__ w) {
  using O = MPW<S>;
  extern __shared__ float lds_mp[];
  const int tid = threadIdx.x;
  float4* U = reinterpret_cast<float4*>(lds_mp);              // [256]
  float* H1 = lds_mp + 4 * MP_THREADS;                         // [256][64]
  float* M = H1 + MP_THREADS * MP_C;                           // [256][64] (forces)
  float* GS = M + MP_THREADS * MP_C;                           // [256][NB*MAXS] (forces)
  float* W2S = GS + MP_THREADS * MP_NB * MP_MAXS;              // [NB][C] (forces)
  __shared__ double red[MP_THREADS / 64];
  unsigned* nbl = a.nbl + (size_t)blockIdx.x * MP_MAXNB * MP_THREADS;
  const MPNNDev& P = a.net;
  // W2 in LDS: read as broadcasts inside the pass-3 edge loop
  if (FORCES)
    for (int k = tid; k < MP_NB * MP_C; k += MP_THREADS) W2S[k] = w[O::W2 + k];

  const int B = a.count ? min(*a.count, a.B) : a.B;
  for (int item = blockIdx.x; item < B; item += gridDim.x) {
    const int cfg = a.cfg_index ? a.cfg_index[item] : item;
    const int N = a.n_atoms[cfg];
    const double L = a.cell[(size_t)cfg * 9];
    if (!(L >= (double)P.rc)) {
      // outside the images enumerated below: no energy (the walk rejects)
      if (tid == 0) a.E[cfg] = __builtin_inf();
      if (FORCES && tid < a.Nmax) {
        double* Fo = a.F + ((size_t)cfg * a.Nmax + tid) * 3;
        Fo[0] = Fo[1] = Fo[2] = 0.0;
      }
      continue;
    }
    const float Lf = (float)L, iLf = (float)(1.0 / L);
    const float rcf = P.rc / Lf;  // cutoff in fractional units
    const float rcf2 = rcf * rcf;
    const bool act = tid < N;
    __syncthreads();  // previous item done with LDS
    if (act) {
      const double* p = a.pos + ((size_t)cfg * a.Nmax + tid) * 3;
      double f0 = p[0] / L, f1 = p[1] / L, f2 = p[2] / L;
      f0 -= floor(f0);
      f1 -= floor(f1);
      f2 -= floor(f2);
      U[tid] = make_float4((float)f0, (float)f1, (float)f2,
                           __int_as_float(a.types[(size_t)cfg * a.Nmax + tid]));
    }
    __syncthreads();
    const float4 ui = act ? U[tid] : make_float4(0.f, 0.f, 0.f, 0.f);
    const int ti = act ? __float_as_int(ui.w) : 0;

    // ---- neighbour list
    int nb = 0;
    if (act) {
      for (int j = 0; j < N; ++j) {
        const float4 uj = U[j];
        float d[3] = {uj.x - ui.x, uj.y - ui.y, uj.z - ui.z};
#pragma unroll
        for (int k = 0; k < 3; ++k) d[k] -= rintf(d[k]);
        const int ax = image_axes(d, rcf);
        // the images of the axes in ax, ascending (subsets of ax)
#pragma unroll 1
        for (int m = 0, go = 1; go; m = (m - ax) & ax, go = m != 0) {
          float e0 = (m & 1) ? d[0] - copysignf(1.0f, d[0]) : d[0];
          float e1 = (m & 2) ? d[1] - copysignf(1.0f, d[1]) : d[1];
          float e2 = (m & 4) ? d[2] - copysignf(1.0f, d[2]) : d[2];
          const float r2 = e0 * e0 + e1 * e1 + e2 * e2;
          const bool self = (j == tid) && m == 0;
          if (r2 < rcf2 && !self) {
            if (nb < MP_MAXNB) nbl[(size_t)nb * MP_THREADS + tid] = (unsigned)j | ((unsigned)m << 9);
            ++nb;
          }
        }
      }
      if (nb > MP_MAXNB) {
        atomicOr(a.overflow, 1);
        nb = MP_MAXNB;
      }
    }

    // ---- pass 1: S and core energy
    float Sacc[MP_NB][MP_MAXS];
#pragma unroll
    for (int n = 0; n < MP_NB; ++n)
#pragma unroll
      for (int s = 0; s < MP_MAXS; ++s) Sacc[n][s] = 0.0f;
    float ecore = 0.0f;
    const float sig = P.core_sig, rcc = P.core_rc;
    const float sh = P.core_sh;
    for (int k = 0; k < nb; ++k) {
      float dx, dy, dz;
      int j, tj;
      mp_edge(nbl[(size_t)k * MP_THREADS + tid], U, ui, dx, dy, dz, j, tj);
      const float d2 = dx * dx + dy * dy + dz * dz;
      const float id = __builtin_amdgcn_rsqf(d2);
      const float r = Lf * d2 * id, ir = id * iLf;
      float b[MP_NB], db[MP_NB];
      mp_basis<false>(r, ir, P, b, db);
#pragma unroll
      for (int n = 0; n < MP_NB; ++n)
#pragma unroll
        for (int s = 0; s < MP_MAXS; ++s)
          if (s == tj) Sacc[n][s] += b[n];
      if (r < rcc) {
        const float q = sig * ir, q2 = q * q, q6 = q2 * q2 * q2;
        ecore += 0.5f * P.core_eps * (q6 * q6 - sh);
      }
    }
    // h1 = emb[t_i] + silu(m1 U1 + c1)
    float m[MP_C];
#pragma unroll
    for (int c = 0; c < MP_C; ++c) {
      float acc = 0.0f;
#pragma unroll
      for (int n = 0; n < MP_NB; ++n) {
        float t = 0.0f;
#pragma unroll
        for (int s = 0; s < S; ++s) t += w[O::emb + s * MP_C + c] * Sacc[n][s];
        acc += w[O::W1 + n * MP_C + c] * t;
      }
      m[c] = acc;
    }
    float z[MP_C];
#pragma unroll
    for (int d = 0; d < MP_C; ++d) z[d] = w[O::c1 + d];
#pragma unroll
    for (int c = 0; c < MP_C; ++c) {
      const float mc = m[c];
#pragma unroll
      for (int d = 0; d < MP_C; ++d) z[d] += mc * w[O::U1 + c * MP_C + d];
    }
    if (act) {
#pragma unroll
      for (int d = 0; d < MP_C; ++d)
        H1[tid * MP_C + d] = w[O::emb + ti * MP_C + d] + silu_f(z[d]);
    }
    __syncthreads();

    // ---- pass 2: A2 chunks -> m2
#pragma unroll
    for (int q = 0; q < MP_C / MP_CH; ++q) {
      float acc[MP_NB][MP_CH];
#pragma unroll
      for (int n = 0; n < MP_NB; ++n)
#pragma unroll
        for (int c = 0; c < MP_CH; ++c) acc[n][c] = 0.0f;
      for (int k = 0; k < nb; ++k) {
        float dx, dy, dz;
        int j, tj;
        mp_edge(nbl[(size_t)k * MP_THREADS + tid], U, ui, dx, dy, dz, j, tj);
        const float d2 = dx * dx + dy * dy + dz * dz;
        const float id = __builtin_amdgcn_rsqf(d2);
        const float r = Lf * d2 * id, ir = id * iLf;
        float b[MP_NB], db[MP_NB];
        mp_basis<false>(r, ir, P, b, db);
        const float4* hj = reinterpret_cast<const float4*>(H1 + j * MP_C + q * MP_CH);
        float h[MP_CH];
#pragma unroll
        for (int v = 0; v < MP_CH / 4; ++v) {
          const float4 t = hj[v];
          h[4 * v] = t.x;
          h[4 * v + 1] = t.y;
          h[4 * v + 2] = t.z;
          h[4 * v + 3] = t.w;
        }
#pragma unroll
        for (int n = 0; n < MP_NB; ++n)
#pragma unroll
          for (int c = 0; c < MP_CH; ++c) acc[n][c] += b[n] * h[c];
      }
#pragma unroll
      for (int c = 0; c < MP_CH; ++c) {
        float t = 0.0f;
#pragma unroll
        for (int n = 0; n < MP_NB; ++n) t += w[O::W2 + n * MP_C + q * MP_CH + c] * acc[n][c];
        m[q * MP_CH + c] = t;
      }
    }
#pragma unroll
    for (int d = 0; d < MP_C; ++d) z[d] = w[O::c2 + d];
#pragma unroll
    for (int c = 0; c < MP_C; ++c) {
      const float mc = m[c];
#pragma unroll
      for (int d = 0; d < MP_C; ++d) z[d] += mc * w[O::U2 + c * MP_C + d];
    }
    float ei = ecore;
    if (act) {
#pragma unroll
      for (int d = 0; d < MP_C; ++d)
        ei += w[O::wo + d] * (H1[tid * MP_C + d] + silu_f(z[d]));
    } else {
      ei = 0.0f;
    }
    // block sum of the per-atom energies in FP64
    double ed = (double)ei;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) ed += __shfl_xor(ed, off, 64);
    if ((tid & 63) == 0) red[tid >> 6] = ed;
    __syncthreads();
    if (tid == 0) {
      double e = 0.0;
      for (int w = 0; w < MP_THREADS / 64; ++w) e += red[w];
      a.E[cfg] = e;
    }
    if (!FORCES) continue;

    // ---- backward: M_i = dE/dm2_i = U2 (wo * silu'(z2))
#pragma unroll
    for (int d = 0; d < MP_C; ++d) z[d] = w[O::wo + d] * silu_d_f(z[d]);
    if (act) {
#pragma unroll
      for (int c = 0; c < MP_C; ++c) {
        float t = 0.0f;
#pragma unroll
        for (int d = 0; d < MP_C; ++d) t += w[O::U2 + c * MP_C + d] * z[d];
        M[tid * MP_C + c] = t;
      }
    }
    __syncthreads();

    // ---- pass 3: layer-2 pair derivatives and dE/dh1
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    float gh[MP_C];
#pragma unroll
    for (int q = 0; q < MP_C / MP_CH; ++q) {
      float qa[MP_NB][MP_CH];
#pragma unroll
      for (int n = 0; n < MP_NB; ++n)
#pragma unroll
        for (int c = 0; c < MP_CH; ++c) qa[n][c] = 0.0f;
      float mi[MP_CH], hi[MP_CH];
#pragma unroll
      for (int c = 0; c < MP_CH; ++c) {
        mi[c] = act ? M[tid * MP_C + q * MP_CH + c] : 0.0f;
        hi[c] = act ? H1[tid * MP_C + q * MP_CH + c] : 0.0f;
      }
      for (int k = 0; k < nb; ++k) {
        float dx, dy, dz;
        int j, tj;
        mp_edge(nbl[(size_t)k * MP_THREADS + tid], U, ui, dx, dy, dz, j, tj);
        const float d2 = dx * dx + dy * dy + dz * dz;
        const float id = __builtin_amdgcn_rsqf(d2);
        const float r = Lf * d2 * id, ir = id * iLf;
        float b[MP_NB], db[MP_NB];
        mp_basis<true>(r, ir, P, b, db);
        const float4* hj = reinterpret_cast<const float4*>(H1 + j * MP_C + q * MP_CH);
        const float4* mj = reinterpret_cast<const float4*>(M + j * MP_C + q * MP_CH);
        float v[MP_CH], mjv[MP_CH];
#pragma unroll
        for (int w = 0; w < MP_CH / 4; ++w) {
          const float4 th = hj[w], tm = mj[w];
          v[4 * w] = mi[4 * w] * th.x + tm.x * hi[4 * w];
          v[4 * w + 1] = mi[4 * w + 1] * th.y + tm.y * hi[4 * w + 1];
          v[4 * w + 2] = mi[4 * w + 2] * th.z + tm.z * hi[4 * w + 2];
          v[4 * w + 3] = mi[4 * w + 3] * th.w + tm.w * hi[4 * w + 3];
          mjv[4 * w] = tm.x;
          mjv[4 * w + 1] = tm.y;
          mjv[4 * w + 2] = tm.z;
          mjv[4 * w + 3] = tm.w;
        }
        float g = 0.0f;
#pragma unroll
        for (int n = 0; n < MP_NB; ++n) {
          float t = 0.0f;
#pragma unroll
          for (int c = 0; c < MP_CH; ++c) {
            t += W2S[n * MP_C + q * MP_CH + c] * v[c];
            qa[n][c] += b[n] * mjv[c];
          }
          g += t * db[n];
        }
        // F_i += g d / r  (d in Cartesian units = L * (dx, dy, dz))
        const float s = g * id;
        fx += s * dx;
        fy += s * dy;
        fz += s * dz;
      }
#pragma unroll
      for (int c = 0; c < MP_CH; ++c) {
        float t = w[O::wo + q * MP_CH + c];
#pragma unroll
        for (int n = 0; n < MP_NB; ++n) t += w[O::W2 + n * MP_C + q * MP_CH + c] * qa[n][c];
        gh[q * MP_CH + c] = t;
      }
    }
    // layer-1 backward: recompute z1 from S, gS = d E / d S
#pragma unroll
    for (int c = 0; c < MP_C; ++c) {
      float acc = 0.0f;
#pragma unroll
      for (int n = 0; n < MP_NB; ++n) {
        float t = 0.0f;
#pragma unroll
        for (int s = 0; s < S; ++s) t += w[O::emb + s * MP_C + c] * Sacc[n][s];
        acc += w[O::W1 + n * MP_C + c] * t;
      }
      m[c] = acc;
    }
#pragma unroll
    for (int d = 0; d < MP_C; ++d) z[d] = w[O::c1 + d];
#pragma unroll
    for (int c = 0; c < MP_C; ++c) {
      const float mc = m[c];
#pragma unroll
      for (int d = 0; d < MP_C; ++d) z[d] += mc * w[O::U1 + c * MP_C + d];
    }
#pragma unroll
    for (int d = 0; d < MP_C; ++d) z[d] = gh[d] * silu_d_f(z[d]);  // gz1
#pragma unroll
    for (int c = 0; c < MP_C; ++c) {
      float t = 0.0f;
#pragma unroll
      for (int d = 0; d < MP_C; ++d) t += w[O::U1 + c * MP_C + d] * z[d];
      m[c] = t;  // gm1
    }
    if (act) {
#pragma unroll
      for (int n = 0; n < MP_NB; ++n)
#pragma unroll
        for (int s = 0; s < MP_MAXS; ++s) {
          float t = 0.0f;
#pragma unroll
          for (int c = 0; c < MP_C; ++c)
            t += m[c] * w[O::W1 + n * MP_C + c] * w[O::emb + (s < S ? s : 0) * MP_C + c];
          GS[tid * (MP_NB * MP_MAXS) + n * MP_MAXS + s] = (s < S) ? t : 0.0f;
        }
    }
    __syncthreads();

    // ---- pass 4: layer-1 and core pair derivatives
    float gsi[MP_NB][MP_MAXS];
#pragma unroll
    for (int n = 0; n < MP_NB; ++n)
#pragma unroll
      for (int s = 0; s < MP_MAXS; ++s)
        gsi[n][s] = act ? GS[tid * (MP_NB * MP_MAXS) + n * MP_MAXS + s] : 0.0f;
    for (int k = 0; k < nb; ++k) {
      float dx, dy, dz;
      int j, tj;
      mp_edge(nbl[(size_t)k * MP_THREADS + tid], U, ui, dx, dy, dz, j, tj);
      const float d2 = dx * dx + dy * dy + dz * dz;
      const float id = __builtin_amdgcn_rsqf(d2);
      const float r = Lf * d2 * id, ir = id * iLf;
      float b[MP_NB], db[MP_NB];
      mp_basis<true>(r, ir, P, b, db);
      const float* gsj = GS + j * (MP_NB * MP_MAXS);
      float g = 0.0f;
#pragma unroll
      for (int n = 0; n < MP_NB; ++n) {
        float own = 0.0f;
#pragma unroll
        for (int s = 0; s < MP_MAXS; ++s)
          if (s == tj) own = gsi[n][s];
        g += (own + gsj[n * MP_MAXS + ti]) * db[n];
      }
      if (r < rcc) {
        const float qq = sig * ir, q2 = qq * qq, q6 = q2 * q2 * q2;
        g += -12.0f * P.core_eps * q6 * q6 * ir;
      }
      const float s = g * id;
      fx += s * dx;
      fy += s * dy;
      fz += s * dz;
    }
    if (act) {
      // F = -dE/dx_i = sum_pairs g d_hat; d = L (dx, dy, dz), d_hat = d / r
      double* Fo = a.F + ((size_t)cfg * a.Nmax + tid) * 3;
      Fo[0] = (double)fx;
      Fo[1] = (double)fy;
      Fo[2] = (double)fz;
    }
  }
}

}  // namespace jr
