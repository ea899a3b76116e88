// Radial-descriptor MLP surrogate on BF16 matrix cores (FP32 accumulate).
//
//   G_i[s][k] = 1/4 sum_{j within rd, t_j = s} exp(-(r - mu_k)^2 / (2 w^2)) u(r/rd)
//   x_i = [G_i[0][0..31], G_i[1][0..31], onehot(t_i), 0 ...]     (80 inputs)
//   h1 = silu(x W1 + b1), h2 = silu(h1 W2 + b2), e_i = h2 . w3 + b3, E = sum e_i
//
// A wave owns 64 rows (atoms).  Products use v_mfma_f32_32x32x16_bf16 with
// A fragments from LDS rows (row l%32, k = 8(l/32)..+7) and B fragments from
// n-major BF16 weight copies in global memory (column l%32, same k); the
// 32x32 result block holds row 8(v/4) + 4(l/32) + v%4, column l%32 in
// element v of lane l.
#pragma once

#include <hip/hip_runtime.h>

#include "jr_images.h"

// FP32 surrogate: multiply-adds contract to FMA (see jr_mpnn.h).
#pragma clang fp contract(fast)

namespace jr {

constexpr int SG_D = 80;    // inputs (64 descriptor + 2 one-hot + padding)
constexpr int SG_DP = 96;   // inputs padded to 32 for the input-gradient product
constexpr int SG_H = 128;   // hidden units
constexpr int SG_K = 32;    // Gaussians per neighbour species
constexpr float SG_RD = 3.0f, SG_W = 0.1f, SG_SCALE = 0.25f;
constexpr int SG_OFF_W1 = 0;
constexpr int SG_OFF_B1 = SG_OFF_W1 + SG_D * SG_H;
constexpr int SG_OFF_W2 = SG_OFF_B1 + SG_H;
constexpr int SG_OFF_B2 = SG_OFF_W2 + SG_H * SG_H;
constexpr int SG_OFF_W3 = SG_OFF_B2 + SG_H;
constexpr int SG_OFF_B3 = SG_OFF_W3 + SG_H;
constexpr int SG_NP = SG_OFF_B3 + 1;
constexpr int SG_NPP = (SG_NP + 255) & ~255;  // padded partial-gradient stride

typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));

struct SurDev {
  const float* P;     // FP32 master parameters [SG_NP]
  const __bf16* W1t;  // [128][80]  W1 transposed (forward B operand)
  const __bf16* W1r;  // [96][128]  W1 rows, rows >= 80 zero (input gradient)
  const __bf16* W2t;  // [128][128] W2 transposed (forward)
  const __bf16* W2r;  // [128][128] W2 rows (gradient to h1)
};

__device__ __forceinline__ int sg_lane() { return (int)(threadIdx.x & 63); }

__device__ __forceinline__ bf8v ld8(const __bf16* p) {
  return *reinterpret_cast<const bf8v*>(p);
}

__device__ __forceinline__ float sg_silu(float z) {
  return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}
__device__ __forceinline__ float sg_silu_d(float z) {
  const float s = __builtin_amdgcn_rcpf(1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// Row of element v of block (rb, .) held by lane l; the column is nb*32 + l%32.
__device__ __forceinline__ int sg_row(int rb, int v, int l) {
  return rb * 32 + 8 * (v >> 2) + 4 * (l >> 5) + (v & 3);
}

// C[64 x 32 NB] += A[64 x K] B[K x 32 NB]: A rows in LDS (stride lda), B given
// n-major (Bt[n][k], stride ldb) in global memory.
template <int K, int NB>
__device__ __forceinline__ void wave_gemm(const __bf16* A, int lda, const __bf16* Bt,
                                          int ldb, f16v (&c)[2][NB]) {
  const int l = sg_lane(), r = l & 31, kh = (l >> 5) * 8;
#pragma unroll
  for (int kb = 0; kb < K / 16; ++kb) {
    const bf8v a0 = ld8(A + r * lda + kb * 16 + kh);
    const bf8v a1 = ld8(A + (32 + r) * lda + kb * 16 + kh);
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      const bf8v b = ld8(Bt + (nb * 32 + r) * ldb + kb * 16 + kh);
      c[0][nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b, c[0][nb], 0, 0, 0);
      c[1][nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b, c[1][nb], 0, 0, 0);
    }
  }
}

template <int RB, int NB>
__device__ __forceinline__ void zero_acc(f16v (&c)[RB][NB]) {
#pragma unroll
  for (int rb = 0; rb < RB; ++rb)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
      for (int v = 0; v < 16; ++v) c[rb][nb][v] = 0.0f;
}

// Wave-local LDS ordering (the wave's own rows only).
__device__ __forceinline__ void sg_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Forward of the wave's 64 rows: X (LDS, stride 80) -> z1, z2 (FP32, kept in
// the result layout), h1 written to H (LDS, stride 128, BF16).  Returns the
// per-row energies (without b3) summed over the lane's columns, reduced over
// the 32 lanes that share rows: e[rb][v] is the energy of row sg_row(rb,v,l).
__device__ __forceinline__ void sur_forward_rows(const SurDev& W, const __bf16* X,
                                                 __bf16* H, f16v (&z1)[2][4],
                                                 f16v (&z2)[2][4], float (&e)[2][16]) {
  const int l = sg_lane();
  zero_acc<2, 4>(z1);
  wave_gemm<SG_D, 4>(X, SG_D, W.W1t, SG_D, z1);
#pragma unroll
  for (int nb = 0; nb < 4; ++nb) {
    const float b = W.P[SG_OFF_B1 + nb * 32 + (l & 31)];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const float z = z1[rb][nb][v] + b;
        z1[rb][nb][v] = z;
        H[sg_row(rb, v, l) * SG_H + nb * 32 + (l & 31)] = (__bf16)sg_silu(z);
      }
  }
  sg_wave_sync();
  zero_acc<2, 4>(z2);
  wave_gemm<SG_H, 4>(H, SG_H, W.W2t, SG_H, z2);
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int v = 0; v < 16; ++v) e[rb][v] = 0.0f;
#pragma unroll
  for (int nb = 0; nb < 4; ++nb) {
    const int col = nb * 32 + (l & 31);
    const float b = W.P[SG_OFF_B2 + col], w3 = W.P[SG_OFF_W3 + col];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const float z = z2[rb][nb][v] + b;
        z2[rb][nb][v] = z;
        e[rb][v] += w3 * sg_silu(z);
      }
  }
  // sum over the 32 lanes holding the same rows (xor over lane bits 0..4)
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      float t = e[rb][v];
#pragma unroll
      for (int m = 1; m < 32; m <<= 1) t += __shfl_xor(t, m, 64);
      e[rb][v] = t;
    }
}

// ---------------------------------------------------------------- eval
// ---------------------------------------------------------------- eval
// Energy (and forces) of a batch of configurations in three passes:
//   sur_desc_kernel   one block per configuration: descriptor rows (BF16,
//                     [B][Nmax][80], zero beyond N) and the pair baseline
//   sur_mlp_kernel    one wave per 32-row tile on the matrix cores: row
//                     energies and, for forces, dE/dG rows ([B][Nmax][64])
//   sur_energy_kernel sum of the row energies, N b3 and the baseline
//   sur_force_kernel  one block per configuration: pair forces from the
//                     dE/dG rows staged in LDS
// Splitting keeps each pass's registers and LDS small enough to run several
// waves per SIMD; the rows travel through HBM (well below a millisecond).
struct SurCfgArgs {
  int B, Nmax;
  const double* pos;     // [B][Nmax][3]
  const int* types;      // [B][Nmax]
  const int* n_atoms;    // [B]
  const double* cell;    // [B][9] cubic
  const int* cfg_index;  // [B] or NULL
  const int* count;      // device item count (<= B), or NULL (= B)
  __bf16* X;             // [B][Nmax][80] descriptor rows
  const float* dX;       // [B][Nmax][64] dE/dG rows (force pass)
  double* F;             // [B][Nmax][3] (force pass)
  double* Ebase;         // [B] baseline energy (descriptor pass)
  // pair baseline b_eps ((sig/r)^12 - shift) for r^2 < b_rc2 (b_eps = 0: none)
  float b_eps, b_sig2, b_rc2, b_shift;
};

__device__ __forceinline__ int sur_n_items(const int* count, int B) {
  return count ? min(*count, B) : B;
}

// Gaussians g_k(r) (and derivatives when D).
// The centres are evenly spaced (delta apart), so with t_k = r - mu_k
//   g_{k+1} = g_k R_k,  R_k = exp((delta t_k - delta^2 / 2) / w^2),
//   R_{k+1} = R_k exp(-delta^2 / w^2):
// each block of 8 centres takes two exponentials (its first Gaussian and
// ratio) instead of eight.  Relative error ~1e-6, far below the BF16 rounding
// of the descriptor; a block whose first Gaussian underflows contributes
// nothing (its values are below 3e-20 there).
template <bool D>
__device__ __forceinline__ void sg_gauss(float r, float (&g)[SG_K], float (&dg)[SG_K]) {
  const float x = r * (1.0f / SG_RD);
  const float x2 = x * x, x5 = x2 * x2 * x, x6 = x5 * x, x7 = x6 * x;
  const float u = 1.0f - 28.0f * x6 + 48.0f * x7 - 21.0f * x7 * x;
  const float du = (-168.0f * x5 + 336.0f * x6 - 168.0f * x7) * (1.0f / SG_RD);
  constexpr float iw2 = 1.0f / (SG_W * SG_W);
  constexpr float delta = 1.7f / 31.0f;
  constexpr float c = -delta * delta * iw2;  // log of R_{k+1} / R_k
#pragma unroll
  for (int blk = 0; blk < SG_K / 8; ++blk) {
    const float t0 = r - (0.8f + 1.7f * (float)(8 * blk) / 31.0f);
    float ex = __expf(-0.5f * t0 * t0 * iw2);
    float R = __expf(delta * t0 * iw2 + 0.5f * c);
    const float q = __expf(c);
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const int k = 8 * blk + m;
      g[k] = ex * u;
      if (D) {
        const float t = r - (0.8f + 1.7f * (float)k / 31.0f);
        dg[k] = ex * (du - t * iw2 * u);
      }
      ex *= R;
      R *= q;
    }
  }
}

// Fractional coordinates (wrapped) and species of one configuration into U.
__device__ __forceinline__ void sur_stage_positions(const SurCfgArgs& a, int cfg, int N,
                                                    double L, float4* U) {
  const int tid = threadIdx.x;
  if (tid < N) {
    const double* p = a.pos + ((size_t)cfg * a.Nmax + tid) * 3;
    double f0 = p[0] / L, f1 = p[1] / L, f2 = p[2] / L;
    f0 -= floor(f0);
    f1 -= floor(f1);
    f2 -= floor(f2);
    U[tid] = make_float4((float)f0, (float)f1, (float)f2,
                         __int_as_float(a.types[(size_t)cfg * a.Nmax + tid]));
  }
}

__global__ void __launch_bounds__(256) sur_desc_kernel(SurCfgArgs a) {
  __shared__ float4 U[256];
  __shared__ double red[4];
  const int item = blockIdx.x;
  if (item >= sur_n_items(a.count, a.B)) return;  // block-uniform
  const int cfg = a.cfg_index ? a.cfg_index[item] : item;
  const int tid = threadIdx.x, l = tid & 63, wv = tid >> 6;
  const int N = a.n_atoms[cfg];
  const double L = a.cell[(size_t)cfg * 9];
  const float Lf = (float)L, iLf = (float)(1.0 / L);
  const float rdf2 = (SG_RD * iLf) * (SG_RD * iLf);
  const float brc2 = a.b_rc2 * iLf * iLf;  // baseline cutoff, fractional units
  const bool act = tid < N;
  sur_stage_positions(a, cfg, N, L, U);
  __syncthreads();
  const float4 ui = act ? U[tid] : make_float4(0.f, 0.f, 0.f, 0.f);
  const int ti = act ? __float_as_int(ui.w) : 0;
  float G0[SG_K], G1[SG_K];
#pragma unroll
  for (int k = 0; k < SG_K; ++k) G0[k] = G1[k] = 0.0f;
  float eb = 0.0f;  // sum over partners of the baseline pair energy
  if (act) {
    for (int j = 0; j < N; ++j) {
      const float4 uj = U[j];
      float d[3] = {uj.x - ui.x, uj.y - ui.y, uj.z - ui.z};
#pragma unroll
      for (int k = 0; k < 3; ++k) d[k] -= rintf(d[k]);
      const int tj = __float_as_int(uj.w);
      const int ax = image_axes(d, SG_RD * iLf);
      // the images of the axes in ax, ascending (subsets of ax)
#pragma unroll 1
      for (int m = 0, go = 1; go; m = (m - ax) & ax, go = m != 0) {
        const float e0 = (m & 1) ? d[0] - copysignf(1.0f, d[0]) : d[0];
        const float e1 = (m & 2) ? d[1] - copysignf(1.0f, d[1]) : d[1];
        const float e2 = (m & 4) ? d[2] - copysignf(1.0f, d[2]) : d[2];
        const float r2 = e0 * e0 + e1 * e1 + e2 * e2;
        if (!(r2 < rdf2) || (j == tid && m == 0)) continue;
        const float r = Lf * sqrtf(r2);
        if (r2 < brc2) {
          const float q = a.b_sig2 * iLf * iLf / r2, q3 = q * q * q;
          eb += a.b_eps * (q3 * q3 - a.b_shift);
        }
        float g[SG_K], dg[SG_K];
        sg_gauss<false>(r, g, dg);
#pragma unroll
        for (int k = 0; k < SG_K; ++k) {
          if (tj == 0) G0[k] += g[k];
          else G1[k] += g[k];
        }
      }
    }
  }
  if (tid < a.Nmax) {
    __bf16 row[SG_D];
#pragma unroll
    for (int k = 0; k < SG_D; ++k) {
      float v = 0.0f;
      if (k < SG_K) v = SG_SCALE * G0[k];
      else if (k < 2 * SG_K) v = SG_SCALE * G1[k - SG_K];
      else if (k == 2 * SG_K + ti) v = 1.0f;
      row[k] = (__bf16)(act ? v : 0.0f);
    }
    bf8v* o = reinterpret_cast<bf8v*>(a.X + ((size_t)cfg * a.Nmax + tid) * SG_D);
#pragma unroll
    for (int q = 0; q < SG_D / 8; ++q) {
      bf8v v;
#pragma unroll
      for (int e = 0; e < 8; ++e) v[e] = row[8 * q + e];
      o[q] = v;
    }
  }
  double edb = act ? 0.5 * (double)eb : 0.0;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) edb += __shfl_xor(edb, off, 64);
  if (l == 0) red[wv] = edb;
  __syncthreads();
  if (tid == 0) a.Ebase[cfg] = (red[0] + red[1]) + (red[2] + red[3]);
}

// C[32 RB x 32 NB] += A[32 RB x K] B[K x 32 NB]: A rows in LDS (stride lda),
// B given n-major (Bt[n][k], stride ldb) in global memory.
template <int K, int NB, int RB>
__device__ __forceinline__ void wave_gemm_rb(const __bf16* A, int lda, const __bf16* Bt, int ldb,
                                             f16v (&c)[RB][NB]) {
  const int l = sg_lane(), r = l & 31, kh = (l >> 5) * 8;
#pragma unroll 2
  for (int kb = 0; kb < K / 16; ++kb) {
    bf8v av[RB];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) av[rb] = ld8(A + (32 * rb + r) * lda + kb * 16 + kh);
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      const bf8v b = ld8(Bt + (nb * 32 + r) * ldb + kb * 16 + kh);
#pragma unroll
      for (int rb = 0; rb < RB; ++rb)
        c[rb][nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[rb], b, c[rb][nb], 0, 0, 0);
    }
  }
}

struct SurMlpArgs {
  SurDev net;
  int B, Nmax, tiles;    // tiles of 32 rows per configuration
  const int* cfg_index;  // [B] or NULL
  const int* count;      // device item count, or NULL
  const __bf16* X;       // [B][Nmax][80]
  float* erow;           // [B][Nmax] row energies (without b3)
  float* dX;             // [B][Nmax][64] dE/dG rows, or NULL (energy only)
};

typedef __bf16 bf2v __attribute__((ext_vector_type(2)));

// One wave per 32-row tile.  silu'(z1) is kept as BF16 pairs (the operand
// rounding the next product applies anyway) and z2 is consumed into the row
// energies and the BF16 tile G2 = w3 silu'(z2), so no pre-activation stays live
// across a matrix product.
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 8)))
    sur_mlp_kernel(SurMlpArgs a) {
  __shared__ __attribute__((aligned(16))) __bf16 Xs[32 * SG_D];
  __shared__ __attribute__((aligned(16))) __bf16 Hs[32 * SG_H];
  const int item = blockIdx.x;
  if (item >= sur_n_items(a.count, a.B) * a.tiles) return;
  const int q = item / a.tiles, tile = item - q * a.tiles;
  const int cfg = a.cfg_index ? a.cfg_index[q] : q;
  const int row0 = tile * 32;
  const int l = sg_lane();
  const size_t base = (size_t)cfg * a.Nmax + row0;
  for (int i = l; i < 32 * SG_D / 8; i += 64) {
    const int r = i / (SG_D / 8), part = i - r * (SG_D / 8);
    bf8v v;
    if (row0 + r < a.Nmax) {
      v = ld8(a.X + (base + r) * SG_D + part * 8);
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) v[e] = (__bf16)0.0f;
    }
    *reinterpret_cast<bf8v*>(Xs + r * SG_D + part * 8) = v;
  }
  sg_wave_sync();
  const bool want_f = a.dX != nullptr;
  bf2v s1[4][8];
  float e[16];
  {
    f16v z[1][4];
    zero_acc<1, 4>(z);
    wave_gemm_rb<SG_D, 4, 1>(Xs, SG_D, a.net.W1t, SG_D, z);
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
      const float b = a.net.P[SG_OFF_B1 + nb * 32 + (l & 31)];
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const float zz = z[0][nb][v] + b;
        const float sg = __builtin_amdgcn_rcpf(1.0f + __expf(-zz));
        Hs[sg_row(0, v, l) * SG_H + nb * 32 + (l & 31)] = (__bf16)(zz * sg);
        s1[nb][v >> 1][v & 1] = (__bf16)(sg * (1.0f + zz * (1.0f - sg)));
      }
    }
  }
  sg_wave_sync();
  {
    f16v z[1][4];
    zero_acc<1, 4>(z);
    wave_gemm_rb<SG_H, 4, 1>(Hs, SG_H, a.net.W2t, SG_H, z);
    sg_wave_sync();  // h1 fully read before G2 replaces it
#pragma unroll
    for (int v = 0; v < 16; ++v) e[v] = 0.0f;
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
      const int col = nb * 32 + (l & 31);
      const float b = a.net.P[SG_OFF_B2 + col], w3 = a.net.P[SG_OFF_W3 + col];
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const float zz = z[0][nb][v] + b;
        const float sg = __builtin_amdgcn_rcpf(1.0f + __expf(-zz));
        e[v] += w3 * (zz * sg);
        if (want_f)
          Hs[sg_row(0, v, l) * SG_H + col] = (__bf16)(w3 * (sg * (1.0f + zz * (1.0f - sg))));
      }
    }
  }
  // row energies: sum over the 32 lanes holding the same rows
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    float t = e[v];
#pragma unroll
    for (int m = 1; m < 32; m <<= 1) t += __shfl_xor(t, m, 64);
    e[v] = t;
  }
  if ((l & 31) == 0) {
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int r = sg_row(0, v, l);
      if (row0 + r < a.Nmax) a.erow[base + r] = e[v];
    }
  }
  if (!want_f) return;
  sg_wave_sync();
  {
    f16v g[1][4];
    zero_acc<1, 4>(g);
    wave_gemm_rb<SG_H, 4, 1>(Hs, SG_H, a.net.W2r, SG_H, g);
    sg_wave_sync();
#pragma unroll
    for (int nb = 0; nb < 4; ++nb)
#pragma unroll
      for (int v = 0; v < 16; ++v)
        Hs[sg_row(0, v, l) * SG_H + nb * 32 + (l & 31)] =
            (__bf16)(g[0][nb][v] * (float)s1[nb][v >> 1][v & 1]);
  }
  sg_wave_sync();
  f16v gx[1][2];
  zero_acc<1, 2>(gx);
  wave_gemm_rb<SG_H, 2, 1>(Hs, SG_H, a.net.W1r, SG_H, gx);
#pragma unroll
  for (int nb = 0; nb < 2; ++nb)
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int r = sg_row(0, v, l);
      if (row0 + r < a.Nmax) a.dX[(base + r) * 64 + nb * 32 + (l & 31)] = gx[0][nb][v];
    }
}

// E = sum of the first N row energies + N b3 + baseline (one wave per
// configuration, fixed order).
__global__ void __launch_bounds__(256) sur_energy_kernel(int B, int Nmax, const int* cfg_index,
                                                         const int* count, const int* n_atoms,
                                                         const float* erow, const float* P,
                                                         const double* Ebase, double* E) {
  const int item = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (item >= sur_n_items(count, B)) return;
  const int cfg = cfg_index ? cfg_index[item] : item;
  const int N = n_atoms[cfg];
  double s = 0.0;
  for (int i = (int)(threadIdx.x & 63); i < N; i += 64) s += (double)erow[(size_t)cfg * Nmax + i];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) E[cfg] = s + (double)N * (double)P[SG_OFF_B3] + Ebase[cfg];
}

// Pair forces: F_i = sum_j (dE/dr_ij) (x_j - x_i)/r_ij with dE/dr from both
// atoms' descriptor gradients (dE/dG rows staged in LDS) plus the baseline.
__global__ void __launch_bounds__(256) sur_force_kernel(SurCfgArgs a) {
  extern __shared__ float lds_f[];
  float4* U = reinterpret_cast<float4*>(lds_f);  // [256]
  float* dX = lds_f + 4 * 256;                    // [Nmax][64]
  const int item = blockIdx.x;
  if (item >= sur_n_items(a.count, a.B)) return;
  const int cfg = a.cfg_index ? a.cfg_index[item] : item;
  const int tid = threadIdx.x;
  const int N = a.n_atoms[cfg];
  const double L = a.cell[(size_t)cfg * 9];
  const float Lf = (float)L, iLf = (float)(1.0 / L);
  const float rdf2 = (SG_RD * iLf) * (SG_RD * iLf);
  const float brc2 = a.b_rc2 * iLf * iLf;
  const bool act = tid < N;
  sur_stage_positions(a, cfg, N, L, U);
  const float* src = a.dX + (size_t)cfg * a.Nmax * 64;
  for (int i = tid; i < N * 64; i += 256) dX[i] = src[i];
  __syncthreads();
  if (act) {
    const float4 ui = U[tid];
    const int ti = __float_as_int(ui.w);
    float own[2 * SG_K];
#pragma unroll
    for (int k = 0; k < 2 * SG_K; ++k) own[k] = dX[tid * 64 + k];
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    for (int j = 0; j < N; ++j) {
      const float4 uj = U[j];
      float d[3] = {uj.x - ui.x, uj.y - ui.y, uj.z - ui.z};
#pragma unroll
      for (int k = 0; k < 3; ++k) d[k] -= rintf(d[k]);
      const int tj = __float_as_int(uj.w);
      const float* other = dX + j * 64 + ti * SG_K;
      const int ax = image_axes(d, SG_RD * iLf);
#pragma unroll 1
      for (int m = 0, go = 1; go; m = (m - ax) & ax, go = m != 0) {
        const float e0 = (m & 1) ? d[0] - copysignf(1.0f, d[0]) : d[0];
        const float e1 = (m & 2) ? d[1] - copysignf(1.0f, d[1]) : d[1];
        const float e2 = (m & 4) ? d[2] - copysignf(1.0f, d[2]) : d[2];
        const float r2 = e0 * e0 + e1 * e1 + e2 * e2;
        if (!(r2 < rdf2) || (j == tid && m == 0)) continue;
        const float id = __builtin_amdgcn_rsqf(r2);
        const float r = Lf * r2 * id;
        float gg[SG_K], dg[SG_K];
        sg_gauss<true>(r, gg, dg);
        float c = 0.0f;
#pragma unroll
        for (int k = 0; k < SG_K; ++k) {
          const float mine = (tj == 0) ? own[k] : own[SG_K + k];
          c += (mine + other[k]) * dg[k];
        }
        c *= SG_SCALE;
        if (r2 < brc2) {  // d phi / dr of the baseline pair
          const float q = a.b_sig2 * iLf * iLf / r2, q3 = q * q * q;
          c -= 12.0f * a.b_eps * q3 * q3 / r;
        }
        const float s = c * id;  // d / r = (e0, e1, e2) * id
        fx += s * e0;
        fy += s * e1;
        fz += s * e2;
      }
    }
    double* Fo = a.F + ((size_t)cfg * a.Nmax + tid) * 3;
    Fo[0] = (double)fx;
    Fo[1] = (double)fy;
    Fo[2] = (double)fz;
  } else if (tid < a.Nmax) {
    double* Fo = a.F + ((size_t)cfg * a.Nmax + tid) * 3;
    Fo[0] = Fo[1] = Fo[2] = 0.0;
  }
}

// ---------------------------------------------------------------- training
// Rows are descriptor rows of exactly evaluated configurations, concatenated;
// a 64-row chunk per one-wave block, rows beyond n_rows padded with zeros.
struct SurRowsArgs {
  SurDev net;
  int n_rows;
  const __bf16* x;     // [n_rows][80]
  float* e;            // [n_rows] row energies (forward), or NULL
  const float* c;      // [n_rows] dLoss/de per row (backward), or NULL
  float* gpart;        // [n_chunks][SG_NPP] partial gradients (backward)
};

__device__ __forceinline__ void load_rows(const SurRowsArgs& a, int chunk, __bf16* X,
                                          __bf16* XT) {
  const int l = sg_lane();
  const int row = chunk * 64 + l;
  const bool ok = row < a.n_rows;
#pragma unroll
  for (int k = 0; k < SG_D; k += 8) {
    bf8v v;
    if (ok) v = ld8(a.x + (size_t)row * SG_D + k);
    else
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = (__bf16)0.0f;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      X[l * SG_D + k + q] = v[q];
      if (XT) XT[(k + q) * 64 + l] = v[q];
    }
  }
  if (XT)
#pragma unroll
    for (int k = SG_D; k < SG_DP; ++k) XT[k * 64 + l] = (__bf16)0.0f;
  sg_wave_sync();
}

__global__ void __launch_bounds__(64) sur_rows_forward_kernel(SurRowsArgs a) {
  __shared__ __attribute__((aligned(16))) __bf16 X[64 * SG_D];
  __shared__ __attribute__((aligned(16))) __bf16 H[64 * SG_H];
  const int l = sg_lane();
  for (int chunk = blockIdx.x; chunk * 64 < a.n_rows; chunk += gridDim.x) {
    load_rows(a, chunk, X, nullptr);
    f16v z1[2][4], z2[2][4];
    float e[2][16];
    sur_forward_rows(a.net, X, H, z1, z2, e);
    const float b3 = a.net.P[SG_OFF_B3];
    if ((l & 31) == 0)
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int row = chunk * 64 + sg_row(rb, v, l);
          if (row < a.n_rows) a.e[row] = e[rb][v] + b3;
        }
    sg_wave_sync();
  }
}

// Weight-gradient block C[32 x 32] (M block mb, N block nb) += sum over the 64
// rows of AT[m][row] BT[n][row]: both operands transposed in LDS (stride 64).
__device__ __forceinline__ f16v tgemm(const __bf16* AT, const __bf16* BT, int mb, int nb) {
  const int l = sg_lane(), r = l & 31, kh = (l >> 5) * 8;
  f16v c;
#pragma unroll
  for (int v = 0; v < 16; ++v) c[v] = 0.0f;
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) {
    const bf8v x = ld8(AT + (mb * 32 + r) * 64 + kb * 16 + kh);
    const bf8v y = ld8(BT + (nb * 32 + r) * 64 + kb * 16 + kh);
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x, y, c, 0, 0, 0);
  }
  return c;
}

__global__ void __launch_bounds__(64) sur_rows_backward_kernel(SurRowsArgs a) {
  __shared__ __attribute__((aligned(16))) __bf16 X[64 * SG_D];
  __shared__ __attribute__((aligned(16))) __bf16 XT[SG_DP * 64];
  __shared__ __attribute__((aligned(16))) __bf16 H[64 * SG_H];   // h1, then g2 rows
  __shared__ __attribute__((aligned(16))) __bf16 HT[SG_H * 64];  // h1^T, then g1^T
  __shared__ __attribute__((aligned(16))) __bf16 GT[SG_H * 64];  // g2^T
  const int l = sg_lane();
  for (int chunk = blockIdx.x; chunk * 64 < a.n_rows; chunk += gridDim.x) {
    float* gp = a.gpart + (size_t)chunk * SG_NPP;
    load_rows(a, chunk, X, XT);
    f16v z1[2][4], z2[2][4];
    float e[2][16];
    sur_forward_rows(a.net, X, H, z1, z2, e);
    // h1^T for dW2
#pragma unroll
    for (int nb = 0; nb < 4; ++nb)
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int row = sg_row(rb, v, l), col = nb * 32 + (l & 31);
          HT[col * 64 + row] = H[row * SG_H + col];
        }
    // g2 = c * w3 * silu'(z2): rows into H (bf16), transposed into GT;
    // bias and w3 gradients from the unrounded values
    float cr[2][16];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int row = chunk * 64 + sg_row(rb, v, l);
        cr[rb][v] = row < a.n_rows ? a.c[row] : 0.0f;
      }
    sg_wave_sync();
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
      const int col = nb * 32 + (l & 31);
      const float w3 = a.net.P[SG_OFF_W3 + col];
      float sb2 = 0.0f, sw3 = 0.0f;
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int row = sg_row(rb, v, l);
          const float g2 = cr[rb][v] * w3 * sg_silu_d(z2[rb][nb][v]);
          sb2 += g2;
          sw3 += cr[rb][v] * sg_silu(z2[rb][nb][v]);
          const __bf16 gb = (__bf16)g2;
          H[row * SG_H + col] = gb;
          GT[col * 64 + row] = gb;
        }
      sb2 += __shfl_xor(sb2, 32, 64);
      sw3 += __shfl_xor(sw3, 32, 64);
      if (l < 32) {
        gp[SG_OFF_B2 + col] = sb2;
        gp[SG_OFF_W3 + col] = sw3;
      }
    }
    if (l == 0) {
      float s = 0.0f;
      for (int k = 0; k < 64; ++k) {
        const int row = chunk * 64 + k;
        if (row < a.n_rows) s += a.c[row];
      }
      gp[SG_OFF_B3] = s;
    }
    sg_wave_sync();
    // dW2[in2][out] = sum_rows h1[row][in2] g2[row][out]
#pragma unroll 1
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) {
        const f16v c = tgemm(HT, GT, mb, nb);
#pragma unroll
        for (int v = 0; v < 16; ++v)
          gp[SG_OFF_W2 + (mb * 32 + sg_row(0, v, l)) * SG_H + nb * 32 + (l & 31)] = c[v];
      }
    // g1 = (g2 W2^T) * silu'(z1)
    f16v g[2][4];
    zero_acc<2, 4>(g);
    wave_gemm<SG_H, 4>(H, SG_H, a.net.W2r, SG_H, g);
    sg_wave_sync();
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
      const int col = nb * 32 + (l & 31);
      float sb1 = 0.0f;
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const float g1 = g[rb][nb][v] * sg_silu_d(z1[rb][nb][v]);
          sb1 += g1;
          HT[col * 64 + sg_row(rb, v, l)] = (__bf16)g1;
        }
      sb1 += __shfl_xor(sb1, 32, 64);
      if (l < 32) gp[SG_OFF_B1 + col] = sb1;
    }
    sg_wave_sync();
    // dW1[in][h] = sum_rows x[row][in] g1[row][h]  (in < 80)
#pragma unroll 1
    for (int mb = 0; mb < 3; ++mb)
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) {
        const f16v c = tgemm(XT, HT, mb, nb);
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int in = mb * 32 + sg_row(0, v, l);
          if (in < SG_D) gp[SG_OFF_W1 + in * SG_H + nb * 32 + (l & 31)] = c[v];
        }
      }
    sg_wave_sync();
  }
}

// grad[k] = sum over chunks (fixed order) of gpart[chunk][k]
__global__ void sur_reduce_kernel(const float* gpart, int n_chunks, float* grad) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= SG_NP) return;
  float s = 0.0f;
  for (int c = 0; c < n_chunks; ++c) s += gpart[(size_t)c * SG_NPP + k];
  grad[k] = s;
}

// Adam (beta1 0.9, beta2 0.999, eps 1e-8) on the FP32 master parameters.
__global__ void sur_adam_kernel(float* P, float* m, float* v, const float* g, float lr,
                                float c1, float c2) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= SG_NP) return;
  const float gk = g[k];
  const float mk = 0.9f * m[k] + 0.1f * gk;
  const float vk = 0.999f * v[k] + (1.0f - 0.999f) * gk * gk;
  m[k] = mk;
  v[k] = vk;
  P[k] = P[k] - lr * (mk / c1) / (sqrtf(vk / c2) + 1e-8f);
}

// BF16 operand copies of W1 and W2 from the FP32 master parameters.
__global__ void sur_cast_kernel(const float* P, __bf16* W1t, __bf16* W1r, __bf16* W2t,
                                __bf16* W2r) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < SG_D * SG_H) {
    const int in = k / SG_H, h = k % SG_H;
    const __bf16 b = (__bf16)P[SG_OFF_W1 + k];
    W1t[h * SG_D + in] = b;
    W1r[in * SG_H + h] = b;
  }
  if (k < (SG_DP - SG_D) * SG_H) W1r[SG_D * SG_H + k] = (__bf16)0.0f;
  if (k < SG_H * SG_H) {
    const int in = k / SG_H, o = k % SG_H;
    const __bf16 b = (__bf16)P[SG_OFF_W2 + k];
    W2t[o * SG_H + in] = b;
    W2r[in * SG_H + o] = b;
  }
}

}  // namespace jr
