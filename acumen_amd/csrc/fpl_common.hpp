// fpl_common.hpp -- device helpers shared by the LDS forward projectors
// (fpl.hip: raw-pixel boxes; fplc.hip: bilinear-cell boxes).
#pragma once

#include "acu_internal.hpp"

namespace acu {

__device__ __forceinline__ int fpl_find_task(const FPLTask* tasks, int ntasks, int gblock) {
  int lo = 0, hi = ntasks - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (tasks[mid].block_begin <= gblock) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ float fpl_wmin(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ float fpl_wmax(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}

// Range of step offsets u = x - m over which a + b*u stays in [lo, hi].
__device__ __forceinline__ void fpl_urange(float a, float b, float lo, float hi, float& u0, float& u1) {
  if (fabsf(b) < 1e-6f) {
    if (a >= lo && a <= hi) { u0 = -1e30f; u1 = 1e30f; } else { u0 = 1e30f; u1 = -1e30f; }
  } else {
    const float p = (lo - a) / b, q = (hi - a) / b;
    u0 = fminf(p, q);
    u1 = fmaxf(p, q);
  }
}

// DPP moves (gfx9 encodings): row_shr:n = 0x110 + n within rows of 16
// lanes, wave_shl:1 = 0x130 (lane i reads lane i + 1); a lane without a
// source keeps its own value.
template <int CTRL>
__device__ __forceinline__ float fpl_dpp(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), CTRL, 0xF, 0xF, false));
}
// min / max over lanes 0..15 (a prefix scan within the first row, read at
// lane 15), as a wave-uniform value
__device__ __forceinline__ float fpl_row16_min(float v) {
  v = fminf(v, fpl_dpp<0x111>(v));
  v = fminf(v, fpl_dpp<0x112>(v));
  v = fminf(v, fpl_dpp<0x114>(v));
  v = fminf(v, fpl_dpp<0x118>(v));
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 15));
}
__device__ __forceinline__ float fpl_row16_max(float v) {
  v = fmaxf(v, fpl_dpp<0x111>(v));
  v = fmaxf(v, fpl_dpp<0x112>(v));
  v = fmaxf(v, fpl_dpp<0x114>(v));
  v = fmaxf(v, fpl_dpp<0x118>(v));
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 15));
}

// Bounding box of a workgroup's samples over steps [k0, k1] (1-based
// pixel origin with two pixels of margin on every side, size in pixels).
// The workgroup is the rays r0..r1 (consecutive, all < N) x the angles
// a0 .. a0 + na - 1 (na <= 16, those < T).  A sample's position
// fma(c, u, m - fl(s*ym)) and fma(s, u, m + fl(c*ym)) is monotone in the
// ray offset ym -- each rounding step is monotone -- so over the rays the
// extremes are those of the first and last ray, and over the steps those
// of k0 and k1: lanes 0..na-1 take one angle each, four corners, and a
// 16-lane DPP reduction gives exactly the box that reducing over every
// (ray, angle, step) gives.  Wave-uniform; no LDS, no barrier.
struct FplBox { int x0, y0, w, h; bool any; };
struct FplBoxGeom {
  float s, c, px0, px1, py0, py1;  // this lane's angle; m - fl(s*ym), m + fl(c*ym) at the end rays
  bool act;
};
__device__ __forceinline__ FplBoxGeom fpl_box_geom(const FPLTask& tk, int a0, int na, int r0, int r1, int lane) {
#pragma clang fp contract(off)
  FplBoxGeom g;
  const int a = a0 + (lane & 15);
  g.act = lane < na && a < tk.T;
  const float2 sc = const_f2(tk.sc, min(a, tk.T - 1));
  g.s = sc.x;
  g.c = sc.y;
  const float m = tk.mid;
  const float ym0 = (float)(r0 + 1) - m, ym1 = (float)(r1 + 1) - m;
  g.px0 = m - g.s * ym0;
  g.px1 = m - g.s * ym1;
  g.py0 = m + g.c * ym0;
  g.py1 = m + g.c * ym1;
  return g;
}
__device__ __forceinline__ FplBox fpl_box(const FplBoxGeom& g, float m, int k0, int k1) {
#pragma clang fp contract(off)
  float mnx = 1e30f, mny = 1e30f, mxx = -1e30f, mxy = -1e30f;
  if (g.act && k0 <= k1) {
    const float ua = (float)k0 - m, ub = (float)k1 - m;
    const float x00 = fmaf(g.c, ua, g.px0), x01 = fmaf(g.c, ub, g.px0);
    const float x10 = fmaf(g.c, ua, g.px1), x11 = fmaf(g.c, ub, g.px1);
    const float y00 = fmaf(g.s, ua, g.py0), y01 = fmaf(g.s, ub, g.py0);
    const float y10 = fmaf(g.s, ua, g.py1), y11 = fmaf(g.s, ub, g.py1);
    mnx = fminf(fminf(x00, x01), fminf(x10, x11));
    mxx = fmaxf(fmaxf(x00, x01), fmaxf(x10, x11));
    mny = fminf(fminf(y00, y01), fminf(y10, y11));
    mxy = fmaxf(fmaxf(y00, y01), fmaxf(y10, y11));
  }
  mnx = fpl_row16_min(mnx);
  mny = fpl_row16_min(mny);
  mxx = fpl_row16_max(mxx);
  mxy = fpl_row16_max(mxy);
  FplBox b;
  b.any = mnx <= mxx;
  b.x0 = (int)floorf(fmaxf(mnx, -1e6f)) - 2;
  b.y0 = (int)floorf(fmaxf(mny, -1e6f)) - 2;
  b.w = (int)floorf(fminf(mxx, 1e6f)) - b.x0 + 4;
  b.h = (int)floorf(fminf(mxy, 1e6f)) - b.y0 + 4;
  return b;
}

}  // namespace acu
