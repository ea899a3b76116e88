// Device-side building blocks for the gfx950 nested-sampling walk.
//
// Wave64 layout: one wavefront owns one walker; ring position p = lane + 64*s
// (s < A slots per lane) holds atom p.  All reductions over atoms use the
// same canonical order: per-lane partial over slots in ascending order, then
// an xor butterfly over m = 1, 2, 4, 8, 16, 32.  Arithmetic is compiled with
// -ffp-contract=off; only the fma() calls written here fuse.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define JR_WAVE 64

namespace jr {

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

__device__ __forceinline__ double dot3(double a0, double a1, double a2,
                                       double b0, double b1, double b2) {
  return fma(a2, b2, fma(a1, b1, a0 * b0));
}

__device__ __forceinline__ double readlane_d(double v, int l) {
  int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}

// LDS pointers with the address space explicit, so that every access is a
// ds_* instruction whatever the compiler can infer about where a pointer came
// from (a generic pointer to LDS is accessed by slower flat instructions).
template <class T>
using lds_ptr = __attribute__((address_space(3))) T*;
template <class T>
__device__ __forceinline__ lds_ptr<T> to_lds(T* p) {
  return (lds_ptr<T>)p;
}

// DPP move of both 32-bit halves (CTRL: quad_perm, row_mirror, ...).
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
  int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// Butterfly sum over the wave, v[l] += v[l ^ m] for m = 1, 2, .., 32; every
// lane ends with the value of lane 0.  IEEE addition is commutative, so after
// level m the value is constant on aligned blocks of 2m lanes and any lane of
// the partner block can stand in for lane l ^ m: quad_perm for m = 1, 2,
// row_half_mirror (7 - i within 8) for m = 4, row_mirror (15 - i within 16)
// for m = 8, and the four row sums r0..r3 read out for m = 16, 32, giving
// (r0 + r1) + (r2 + r3).  Needs the whole wave active.
__device__ __forceinline__ double wave_sum(double v) {
  v = v + dpp_d<0xB1>(v);   // quad_perm [1,0,3,2]
  v = v + dpp_d<0x4E>(v);   // quad_perm [2,3,0,1]
  v = v + dpp_d<0x141>(v);  // row_half_mirror
  v = v + dpp_d<0x140>(v);  // row_mirror
  const double r0 = readlane_d(v, 0), r1 = readlane_d(v, 16);
  const double r2 = readlane_d(v, 32), r3 = readlane_d(v, 48);
  return (r0 + r1) + (r2 + r3);
}

__device__ __forceinline__ double uniform_d(double v) {
  int lo = __builtin_amdgcn_readfirstlane(__double2loint(v));
  int hi = __builtin_amdgcn_readfirstlane(__double2hiint(v));
  return __hiloint2double(hi, lo);
}

// dst[lane] = src[(lane + 1) % 64]  (DPP wave_rol:1, a VALU move).
__device__ __forceinline__ double wave_rol1(double v) {
  int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x134, 0xF, 0xF, false);
  int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x134, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// One atom per lane, N < 64: new[p] = old[(p + 1) mod N] through the LDS
// crossbar; addr = 4 ((p + 1) mod N) for the active lanes (any lane for the
// idle ones, which then never feed an active position).  Every lane of the
// wave must be active.
__device__ __forceinline__ double ring_perm(double v, int addr) {
  int lo = __builtin_amdgcn_ds_bpermute(addr, __double2loint(v));
  int hi = __builtin_amdgcn_ds_bpermute(addr, __double2hiint(v));
  return __hiloint2double(hi, lo);
}

// Ring rotation with slot s of a lane at ring position A lane + s:
// new[s] = old[s + 1] (registers only) and slot A - 1 takes slot 0 of the
// next lane.  FULL_ (N == 64 A): one DPP wave_rol, whose lane 63 takes
// lane 0.  Otherwise slot 0 moves through the crossbar at byte address addr
// (lane + 1, or lane 0 into the lane holding position N - 1), and the slot
// below A - 1 holding position N - 1 (wrap[s]) takes that value too.
template <int A, bool FULL_>
__device__ __forceinline__ void ring_rot1_ilv(double (&v)[A], int addr,
                                              const bool (&wrap)[A]) {
  const double b = FULL_ ? wave_rol1(v[0]) : ring_perm(v[0], addr);
#pragma unroll
  for (int s = 0; s + 1 < A; ++s) v[s] = (!FULL_ && wrap[s]) ? b : v[s + 1];
  v[A - 1] = b;
}

// Ring rotation when N == 64*A (every ring position occupied): DPP rotate
// each slot, lane 63 takes lane 0 of the next slot.
template <int A>
__device__ __forceinline__ void ring_rot1_full(double (&v)[A], int lane) {
  double r[A];
#pragma unroll
  for (int s = 0; s < A; ++s) r[s] = wave_rol1(v[s]);
#pragma unroll
  for (int s = 0; s < A; ++s) {
    if (A > 1)
      v[s] = (lane == 63) ? r[(s + 1) % A] : r[s];
    else
      v[s] = r[s];
  }
}

// Ring rotation over positions p < N: new[p] = old[(p + 1) % N].  The wave
// rotation carries lane 63 of a slot into lane 0's value of the next slot;
// when the ring is shorter than 64 A (N = 64 under A = 2 included) position
// N - 1 then takes position 0's value instead.
template <int A>
__device__ __forceinline__ void ring_rot1(double (&v)[A], int lane, int N) {
  const bool partial = N != 64 * A;
  double w0 = 0.0;
  if (partial) w0 = readlane_d(v[0], 0);
  double r[A];
#pragma unroll
  for (int s = 0; s < A; ++s) r[s] = wave_rol1(v[s]);
#pragma unroll
  for (int s = 0; s < A; ++s) {
    if (A > 1)
      v[s] = (lane == 63) ? r[(s + 1) % A] : r[s];
    else
      v[s] = r[s];
  }
  if (partial) {
    const int ln = (N - 1) & 63, sn = (N - 1) >> 6;
#pragma unroll
    for (int s = 0; s < A; ++s)
      if (s == sn && lane == ln) v[s] = w0;
  }
}

// Orders LDS traffic of one wavefront (LDS ops of a wave complete in order;
// this keeps the compiler from moving them across the point).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------- cell math
__device__ __forceinline__ double det3(const double* c) {
  double m0 = c[4] * c[8] - c[5] * c[7];
  double m1 = c[3] * c[8] - c[5] * c[6];
  double m2 = c[3] * c[7] - c[4] * c[6];
  return c[0] * m0 - c[1] * m1 + c[2] * m2;
}

__device__ __forceinline__ void inv3(const double* c, double det,
                                     double* inv) {
  double idet = 1.0 / det;
  inv[0] = (c[4] * c[8] - c[5] * c[7]) * idet;
  inv[1] = (c[2] * c[7] - c[1] * c[8]) * idet;
  inv[2] = (c[1] * c[5] - c[2] * c[4]) * idet;
  inv[3] = (c[5] * c[6] - c[3] * c[8]) * idet;
  inv[4] = (c[0] * c[8] - c[2] * c[6]) * idet;
  inv[5] = (c[2] * c[3] - c[0] * c[5]) * idet;
  inv[6] = (c[3] * c[7] - c[4] * c[6]) * idet;
  inv[7] = (c[1] * c[6] - c[0] * c[7]) * idet;
  inv[8] = (c[0] * c[4] - c[1] * c[3]) * idet;
}

// Cube root: integer guess on the IEEE bits, five Newton steps
// y <- (2y + x/y^2)/3.  Deterministic (IEEE +,*,/ only).
__device__ __forceinline__ double cbrt_c(double x) {
  if (!(x > 0.0) || __builtin_isinf(x)) return x;
  const long long bias = (long long)1023 << 52;
  long long b = __double_as_longlong(x);
  long long g = (b - bias) / 3 + bias;
  double y = __longlong_as_double(g);
#pragma unroll
  for (int it = 0; it < 5; ++it) {
    double t = y * y;
    y = (2.0 * y + x / t) / 3.0;
  }
  return y;
}

__device__ __forceinline__ double ipow_c(double x, int n) {
  double r = 1.0, b = x;
  while (n > 0) {
    if (n & 1) r = r * b;
    b = b * b;
    n >>= 1;
  }
  return r;
}

// utils/cell.py:80-125 (reference): volume-per-atom band and min aspect.
__device__ __forceinline__ bool check_cell_shape(const double* c, int n_atoms,
                                                 double max_vpa,
                                                 double min_vpa,
                                                 double min_aspect) {
  double vol = fabs(det3(c));
  double vpa = vol / (double)n_atoms;
  double vc = cbrt_c(vol);
  const int P0[3] = {0, 1, 0}, P1[3] = {1, 2, 2}, P2[3] = {2, 0, 1};
  double aspect = 0.0;
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    const double* a = c + 3 * P0[p];
    const double* b = c + 3 * P1[p];
    const double* k = c + 3 * P2[p];
    double cx = a[1] * b[2] - a[2] * b[1];
    double cy = a[2] * b[0] - a[0] * b[2];
    double cz = a[0] * b[1] - a[1] * b[0];
    double nrm = sqrt(dot3(cx, cy, cz, cx, cy, cz));
    double h = fabs(dot3(cx / nrm, cy / nrm, cz / nrm, k[0], k[1], k[2]));
    double ratio = h / vc;
    aspect = (p == 0) ? ratio : fmin(aspect, ratio);
  }
  return (vpa <= max_vpa) && (vpa >= min_vpa) && (aspect >= min_aspect);
}

// exp(x) from IEEE +, *, rint and exponent arithmetic only, so that every
// build of it rounds alike (Cody and Waite's reduction): x = k ln2 + r with
// ln2 split into a head of 32 significant bits (k * head is exact) and a
// tail, |r| <= ln2 / 2; exp(r) by its Taylor polynomial of degree 13 in
// Horner form (truncation below 5e-18 relative); the result scaled by 2^k in
// two exact power-of-two factors.
__device__ __forceinline__ double pow2i_c(int k) {
  return __longlong_as_double((long long)(k + 1023) << 52);
}

__device__ __forceinline__ double exp_c(double x_in) {
  // out-of-range and NaN inputs take the polynomial path on a clamped value
  // and are replaced by selects at the end (no branches)
  const double x = fmin(fmax(x_in, -0x1.74910d52d3052p+9), 0x1.62e42fefa39efp+9);
  const double k = rint(x * 0x1.71547652b82fep+0);
  const double r = (x - k * 0x1.62e42fee00000p-1) - k * 0x1.a39ef35793c76p-33;
  double p = 0x1.6124613a86d09p-33;
  p = p * r + 0x1.1eed8eff8d898p-29;
  p = p * r + 0x1.ae64567f544e4p-26;
  p = p * r + 0x1.27e4fb7789f5cp-22;
  p = p * r + 0x1.71de3a556c734p-19;
  p = p * r + 0x1.a01a01a01a01ap-16;
  p = p * r + 0x1.a01a01a01a01ap-13;
  p = p * r + 0x1.6c16c16c16c17p-10;
  p = p * r + 0x1.1111111111111p-7;
  p = p * r + 0x1.5555555555555p-5;
  p = p * r + 0x1.5555555555555p-3;
  p = p * r + 0.5;
  p = p * r + 1.0;
  p = p * r + 1.0;
  const int ki = (int)k;
  const int k1 = ki / 2;
  const double y = (p * pow2i_c(k1)) * pow2i_c(ki - k1);
  return x_in != x_in ? x_in
         : x_in > 0x1.62e42fefa39efp+9 ? __builtin_inf()
         : x_in < -0x1.74910d52d3052p+9 ? 0.0
         : y;
}

// sampling/moves/shear.py:19-52: cell vector idx moves by rv1 u + rv2 w,
// u the unit vector along the first other cell vector and w the unit
// vector of the second one's part orthogonal to u.  Rows are picked by
// selects so that the cell stays in registers.
__device__ __forceinline__ void shear_cell(const double* c, int idx,
                                           double rv1, double rv2,
                                           double* nc) {
  double v1[3], v2[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    v1[k] = idx == 0 ? c[3 + k] : c[k];
    v2[k] = idx == 2 ? c[3 + k] : c[6 + k];
  }
  const double n1 = sqrt(dot3(v1[0], v1[1], v1[2], v1[0], v1[1], v1[2]));
  const double u0 = v1[0] / n1, u1 = v1[1] / n1, u2 = v1[2] / n1;
  const double dp = dot3(v2[0], v2[1], v2[2], u0, u1, u2);
  const double w0 = v2[0] - dp * u0, w1 = v2[1] - dp * u1,
               w2 = v2[2] - dp * u2;
  const double n2 = sqrt(dot3(w0, w1, w2, w0, w1, w2));
  const double uu[3] = {u0, u1, u2};
  const double ww[3] = {w0 / n2, w1 / n2, w2 / n2};
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double moved = (c[3 * r + k] + rv1 * uu[k]) + rv2 * ww[k];
      nc[3 * r + k] = r == idx ? moved : c[3 * r + k];
    }
}

// utils/cell.py:128-152: positions follow a cell change as x' = x T with
// T = old^-1 new (rows are lattice vectors).
__device__ __forceinline__ void cell_transform(const double* oc,
                                               const double* nc, double* T) {
  double inv[9];
  inv3(oc, det3(oc), inv);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k)
      T[3 * i + k] = dot3(inv[3 * i], inv[3 * i + 1], inv[3 * i + 2], nc[k],
                          nc[3 + k], nc[6 + k]);
}

// ----------------------------------------------------------- Philox4x32-10
// Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
// SC'11.  Counter-based: (counter, key) -> 4 x 32 bits.
__host__ __device__ __forceinline__ void philox4x32_10(uint32_t c[4],
                                                       uint32_t k0,
                                                       uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += W0;
      k1 += W1;
    }
    uint64_t p0 = (uint64_t)M0 * c[0];
    uint64_t p1 = (uint64_t)M1 * c[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c[1] ^ k0;
    uint32_t n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
  }
}

// 53-bit uniform in [0, 1) from two 32-bit words.
__host__ __device__ __forceinline__ double u53(uint32_t hi, uint32_t lo) {
  uint64_t v = (((uint64_t)hi << 32) | (uint64_t)lo) >> 11;
  return (double)v * 0x1.0p-53;
}

// Philox block `q` of a draw stream: counter (q lo, q hi, it, stream) under
// the key (seed lo, seed hi).
__device__ __forceinline__ void philox_block(unsigned long long seed, long long it,
                                             unsigned stream, long long q, uint32_t c[4]) {
  c[0] = (uint32_t)q;
  c[1] = (uint32_t)((unsigned long long)q >> 32);
  c[2] = (uint32_t)it;
  c[3] = stream;
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// Standard normals, four per Philox block: normals 4q .. 4q+3 of a stream
// come from block q by the Box-Muller transform of its word pairs
// (w0, w1) -> normals 4q, 4q+1 and (w2, w3) -> 4q+2, 4q+3:
//   u = (w + 1/2) 2^-32 in (0, 1],  r = sqrt(-2 ln u_a),
//   z_even = r cos(2 pi u_b),  z_odd = r sin(2 pi u_b),
// in single precision with the hardware transcendentals (v_log_f32 is log2,
// v_sin_f32 / v_cos_f32 take the angle in revolutions), held as FP64 values.
// |z| <= sqrt(2 ln 2 * 33) = 6.76; u rounding to 1 gives z = 0.  These are
// proposal draws; every energy and acceptance test stays FP64.
__device__ __forceinline__ void box_muller(uint32_t wa, uint32_t wb, double& z_cos,
                                           double& z_sin) {
  const float ua = fmaf((float)wa, 0x1.0p-32f, 0x1.0p-33f);
  const float ub = fmaf((float)wb, 0x1.0p-32f, 0x1.0p-33f);
  const float rad = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(ua));
  z_cos = (double)(rad * __builtin_amdgcn_cosf(ub));
  z_sin = (double)(rad * __builtin_amdgcn_sinf(ub));
}

__device__ __forceinline__ void normal_quad(unsigned long long seed, long long it,
                                            unsigned stream, long long q, double z[4]) {
  uint32_t c[4];
  philox_block(seed, it, stream, q, c);
  box_muller(c[0], c[1], z[0], z[1]);
  box_muller(c[2], c[3], z[2], z[3]);
}

// The single normal e of a stream (one Box-Muller pair of block e / 4).
__device__ __forceinline__ double normal_at(unsigned long long seed, long long it,
                                            unsigned stream, long long e) {
  uint32_t c[4];
  philox_block(seed, it, stream, e >> 2, c);
  const bool hi = (e & 2) != 0;
  double zc, zs;
  box_muller(hi ? c[2] : c[0], hi ? c[3] : c[1], zc, zs);
  return (e & 1) ? zs : zc;
}

// Uniform e of a stream: block e / 2 gives u53(w0, w1) for even e and
// u53(w2, w3) for odd e.
__device__ __forceinline__ double uniform_at(unsigned long long seed, long long it,
                                             unsigned stream, long long e) {
  uint32_t c[4];
  philox_block(seed, it, stream, e >> 1, c);
  return (e & 1) ? u53(c[2], c[3]) : u53(c[0], c[1]);
}

}  // namespace jr
