// imdg_device.h -- device helpers shared by the rebuild kernels
// (imdg_build.hip) and the per-step kernels (imdg_step.hip): wavefront and
// compensated reductions, IMD's 3-point table access on knot pairs, the
// grouped neighbour-list layout, cell geometry.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#include "imdg_internal.h"

namespace imdg {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, IMDG_WAVE);
  return v;
}

// Minimum over the 64 lanes (every lane active) through DPP
// moves (quad permutes, row mirrors, row broadcasts of lanes 15 and 31:
// VALU operations without a trip through LDS), read from lane 63.
#define IMDG_DPP_REDUCE(OP, v)                                                \
  do {                                                                        \
    v = OP(v, __builtin_amdgcn_update_dpp(v, v, 0xb1, 0xf, 0xf, false));      \
    v = OP(v, __builtin_amdgcn_update_dpp(v, v, 0x4e, 0xf, 0xf, false));      \
    v = OP(v, __builtin_amdgcn_update_dpp(v, v, 0x141, 0xf, 0xf, false));     \
    v = OP(v, __builtin_amdgcn_update_dpp(v, v, 0x140, 0xf, 0xf, false));     \
    v = OP(v, __builtin_amdgcn_update_dpp(v, v, 0x142, 0xa, 0xf, false));     \
    v = OP(v, __builtin_amdgcn_update_dpp(v, v, 0x143, 0xc, 0xf, false));     \
  } while (0)

__device__ __forceinline__ int wave_min_i(int v) {
  IMDG_DPP_REDUCE(min, v);
  return __builtin_amdgcn_readlane(v, 63);
}


__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, IMDG_WAVE));
  return v;
}

// Error-free sum of two doubles (Knuth TwoSum).
__device__ __forceinline__ void two_sum(double a, double b, double& s,
                                        double& e) {
#pragma clang fp contract(off)
  s = a + b;
  double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// Neumaier-compensated accumulation.
__device__ __forceinline__ void kahan_add(double& s, double& c, double x) {
#pragma clang fp contract(off)
  double t = s + x;
  if (fabs(s) >= fabs(x))
    c += (s - t) + x;
  else
    c += (x - t) + s;
  s = t;
}

__device__ __forceinline__ float rmin(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ double rmin(double a, double b) { return fmin(a, b); }
__device__ __forceinline__ float rmax0(float a) { return fmaxf(a, 0.0f); }
__device__ __forceinline__ double rmax0(double a) { return fmax(a, 0.0); }
__device__ __forceinline__ float rfract(float t) {
  return __builtin_amdgcn_fractf(t);
}
__device__ __forceinline__ double rfract(double t) {
  return __builtin_amdgcn_fract(t);
}

// Segment index and fraction of IMD's 3-point table access
// (src/potaccess.h:329-340): argument clamped to [begin, end], segment
// k = floor((x - begin) / step) capped at len - 1, fraction
// chi = t - floor(t) before the cap (t >= 0, so floor(t) = (int)t; the
// subtraction is exact).  A NaN argument clamps to segment 0.
template <typename R>
__device__ __forceinline__ int seg_locate(R x, R begin, R end, R invstep,
                                          int len, R& chi) {
  R t = rmax0(rmin(x, end) - begin) * invstep;
  int k = (int)t;
  chi = rfract(t);
  return k < len - 1 ? k : len - 1;
}

// IMD's 3-point interpolation on segment k from the knots (p0_k, dv_k) and
// (p0_k+1, dv_k+1), with dv_k = t[k+1] - t[k] and the second difference
// d2v = dv_k+1 - dv_k = t[k+2] - 2 t[k+1] + t[k] (src/potaccess.h:345-353):
// value p0 + chi (dv + (chi - 1)/2 d2v), and twice the derivative with
// respect to the tabulated argument, 2 istep (dv + (chi - 1/2) d2v).
template <typename R, typename R2>
__device__ __forceinline__ R knot_val(R2 a, R2 b, R chi) {
  R d2v = b.y - a.y;
  return a.x + chi * (a.y + (R)0.5 * (chi - (R)1) * d2v);
}
template <typename R, typename R2>
__device__ __forceinline__ R knot_grad(R2 a, R2 b, R chi, R twoistep) {
  R d2v = b.y - a.y;
  return twoistep * (a.y + (chi - (R)0.5) * d2v);
}

// The 4-point and spline interpolations (src/potaccess.h:365-457) are cubic
// in the fraction x of segment k; their tables hold two knots per segment,
// (c0, c1) and (c2, c3): value c0 + x (c1 + x (c2 + x c3)), twice the
// derivative 2 istep (c1 + x (2 c2 + 3 x c3)).  tab_* pick the segment's
// first knot and evaluate it in either form (CUB: cubic segments; else the
// 3-point knots k, k+1).
template <typename R, typename R2>
__device__ __forceinline__ R cub_val(R2 a, R2 b, R x) {
  return a.x + x * (a.y + x * (b.x + x * b.y));
}
template <typename R, typename R2>
__device__ __forceinline__ R cub_grad(R2 a, R2 b, R x, R twoistep) {
  return twoistep * (a.y + x * ((R)2 * b.x + (R)3 * x * b.y));
}
template <bool CUB>
__device__ __forceinline__ int tab_knot(int k) {
  return CUB ? 2 * k : k;
}
template <bool CUB, typename R, typename R2>
__device__ __forceinline__ R tab_val(R2 a, R2 b, R x) {
  if (CUB) return cub_val<R>(a, b, x);
  return knot_val<R>(a, b, x);
}
template <bool CUB, typename R, typename R2>
__device__ __forceinline__ R tab_grad(R2 a, R2 b, R x, R twoistep) {
  if (CUB) return cub_grad<R>(a, b, x, twoistep);
  return knot_grad<R>(a, b, x, twoistep);
}
// Entry m of slot i in the grouped list layout (4 entries per lane).
__device__ __forceinline__ size_t nbr_index(int i, int m, int kmax) {
  return (((size_t)(i >> 6) * (kmax >> 2) + (m >> 2)) * IMDG_WAVE + (i & 63)) *
             IMDG_GROUP +
         (m & 3);
}

// First uint4 of slot i's list; group g is at + g * IMDG_WAVE.
__device__ __forceinline__ const uint4* list_of(const uint32_t* nbr, int i,
                                                int kmax) {
  return reinterpret_cast<const uint4*>(nbr) +
         ((size_t)(i >> 6) * (kmax >> 2)) * IMDG_WAVE + (i & 63);
}

// Byte address of slot i's first 16-bit entry (group 0) in the 16-bit
// format; group g is at + g * 512 bytes, entry m % 4 at + 2 (m % 4).
__device__ __forceinline__ const char* list16_of(const uint32_t* nbr, int i,
                                                 int kmax) {
  return reinterpret_cast<const char*>(nbr) + (size_t)(i >> 6) * kmax * 256 +
         (size_t)(i & 63) * 8;
}

__device__ __forceinline__ void cell_coords(const RepDev& R, int cell,
                                            int& cx, int& cy, int& cz) {
  int lc = cell - R.cell_off;
  cz = lc % R.gdim[2];
  int q = lc / R.gdim[2];
  cy = q % R.gdim[1];
  cx = q / R.gdim[1];
}

// Centre of a cell: sum_d (c_d + 1/2) box_d / gdim_d.
__device__ __forceinline__ void cell_centre(const RepDev& R, int cell,
                                            double* c) {
  int cx, cy, cz;
  cell_coords(R, cell, cx, cy, cz);
  double fx = cx + 0.5, fy = cy + 0.5, fz = cz + 0.5;
#pragma unroll
  for (int e = 0; e < 3; e++)
    c[e] = fx * R.cvec[0][e] + fy * R.cvec[1][e] + fz * R.cvec[2][e];
}

// Squared displacement since the last list build (check_nblist's test,
// src/imd_forces_nbl.c:2016-2024), FP64 with separate roundings.
__device__ __forceinline__ double disp2(double4 a, double4 b) {
#pragma clang fp contract(off)
  double d0 = a.x - b.x, d1 = a.y - b.y, d2 = a.z - b.z;
  return d0 * d0 + d1 * d1 + d2 * d2;
}

// Make the LDS stores of this wavefront visible to its own later loads
// (no block barrier: the waves of a block run independent schedules).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace imdg
