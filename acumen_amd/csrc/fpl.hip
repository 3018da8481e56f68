// fpl.hip -- forward projection (radon, tomography/tv_hpc.F90:129-159) with
// the image staged in LDS and shared by a block of neighbouring angles.
//
// Same definition as fp.hip:
//   sino(y,t) = (1/N) * sum_{x=1..N} bilinear(img; xr, yr), x in order,
//   xr = cos(c)*(x-m) - sin(c)*(y-m) + m,  yr = sin(c)*(x-m) + cos(c)*(y-m) + m,
//   a sample counts only if its floor cell (x1, y1) has 2 <= x1, y1 <= N-1.
//
// gfx950 mapping.  One workgroup = 64 neighbouring rays x a block of 4*al
// neighbouring angles; a lane takes one ray and al angles, a wave 16 rays x
// 4 consecutive angles per instruction (angles g, g+4, ... of the block for
// lane group g), so a half-wave's LDS reads are 16 rays at two adjacent
// angles, whose taps often share dwords, instead of 32 rays at one angle.
// Every lane sums its rays' samples in step order, so the per-ray summation
// order is the reference's.  The sweep goes in chunks of FPL_U steps; for a
// chunk the workgroup stages in LDS the bounding box of all its samples
// (raw FP32 pixels, zero outside the image, a row pitch per angle block
// from a few odd pitches, so both rows of a bilinear cell come from
// ds_read2 immediates and the rays' taps spread over the banks) and every
// angle of the block reads it, so one staged pixel serves ~2-3 samples and
// the gathers never leave the CU (the global-gather kernel of fp.hip moves
// ~23 GB per 2049^2 x 1500 launch; this one stages ~1.7 bytes per sample
// from L2).  The host orders the workgroups so that each XCD gets about the
// same work (fpl_dispatch_order).
//
//  * interior chunks (every cell of the box passes the bounds test):
//    positions are FMAs from the ray's start, there is no test, and a sample
//    is 2 FMA + 2 cvt + 2 fract + addressing + 2 ds_read2_b32 + 3 lerps,
//    the steps software-pipelined over two groups of the lane's angles;
//  * border chunks: positions are the reference's FP32 values (c*u, s*u,
//    s*ym, c*ym rounded once, summed in its order, no contraction) and the
//    bounds test is applied, so every edge decision is the reference's own;
//  * a chunk whose box does not fit the LDS budget falls back to gathering
//    the bilinear cells of fp.hip from global memory (exact positions).
#include "fpl_common.hpp"

#include <algorithm>
#include <cmath>

namespace acu {
namespace {

constexpr int FPL_THREADS = 256;
#ifndef ACUMEN_FPL_AMAX
#define ACUMEN_FPL_AMAX 2
#endif
constexpr int AMAX = ACUMEN_FPL_AMAX;  // max angles per lane (1, 2 or 4)
constexpr int FPL_SB = 8;  // staging loads in flight per lane
constexpr int FPL_SLOTS = 1024;  // workgroups resident at once: 4 per CU x 256 CUs

// Byte offset 4*(iy*PITCH + ix) of the box cell of a position (xl, yl) with
// 0 <= xl, yl < 2^16, ix = (int)xl, iy = (int)yl, in three VALU instructions
// instead of four: ix by a plain conversion, iy converted straight into the
// high half of the same register (SDWA dst_sel:WORD_1, the low half kept),
// then one v_dot2_u32_u16 of the packed pair with (4, 4*PITCH), whose result
// is the LDS byte offset itself.
template <int PITCH>
__device__ __forceinline__ unsigned fpl_cell_offset(float xl, float yl) {
  static_assert(4 * PITCH < 65536, "u16 weight");
  typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
  unsigned pk = (unsigned)(int)xl;
  asm("v_cvt_i32_f32_sdwa %0, %1 dst_sel:WORD_1 dst_unused:UNUSED_PRESERVE src0_sel:DWORD" : "+v"(pk) : "v"(yl));
  u16x2 w;
  w.x = 4;
  w.y = (unsigned short)(4 * PITCH);
  return __builtin_amdgcn_udot2(__builtin_bit_cast(u16x2, pk), w, 0u, false);
}

// One workgroup's work; PITCH is the LDS row pitch of its boxes (which of
// a few pitches spreads the lanes' taps best over the banks depends on the
// angles, see fpl_pitch_classes).
template <int AL, int PITCH>
__device__ __forceinline__ void fpl_body(const FPLTask& tk, int gb, float* __restrict__ box, float (*red)[4]) {
#pragma clang fp contract(off)
  static_assert(PITCH + 1 <= 255, "ds_read2 offsets");
  const int N = tk.N, T = tk.T;
  const float m = tk.mid;
  const int local = gb - tk.block_begin;
  const int rb = local / tk.nab, ab = local - rb * tk.nab;  // ray-block-major
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // wave w: rays 16w..16w+15 (lane & 15) of the block; lane group g = lane >> 4
  // takes angles g, g + 4, ... (so the 64 lanes of one instruction cover 16 rays
  // at 4 consecutive angles)
  const int ray = rb * 64 + wave * 16 + (lane & 15);
  const float ym = (float)(min(ray, N - 1) + 1) - m;

  float s[AL], c[AL], sy[AL], cyv[AL], acc[AL];
  bool act[AL];
  int ang[AL];
  float lu0 = 1e30f, lu1 = -1e30f;
#pragma unroll
  for (int j = 0; j < AL; ++j) {
    ang[j] = ab * 4 * AL + (lane >> 4) + 4 * j;
    act[j] = ang[j] < T && ray < N;
    const float2 v = const_f2(tk.sc, min(ang[j], T - 1));
    s[j] = v.x;
    c[j] = v.y;
    sy[j] = s[j] * ym;   // one rounding each (tv_hpc.F90:143-144)
    cyv[j] = c[j] * ym;
    acc[j] = 0.0f;
    if (act[j]) {
      float ua, ub, va, vb;
      fpl_urange(m - sy[j], c[j], 1.99f, (float)N + 0.01f, ua, ub);
      fpl_urange(m + cyv[j], s[j], 1.99f, (float)N + 0.01f, va, vb);
      const float u0 = fmaxf(ua, va), u1 = fminf(ub, vb);
      if (u0 <= u1) { lu0 = fminf(lu0, u0); lu1 = fmaxf(lu1, u1); }
    }
  }
  // workgroup step window (1-based steps, one step of margin)
  lu0 = fpl_wmin(lu0);
  lu1 = fpl_wmax(lu1);
  if (lane == 0) { red[wave][0] = lu0; red[wave][1] = lu1; }
  if (threadIdx.x == 0) { red[0][2] = 0.0f; red[0][3] = 0.0f; }  // nonzero-box flags of the chunks
  __syncthreads();
  const float U0 = fminf(fminf(red[0][0], red[1][0]), fminf(red[2][0], red[3][0]));
  const float U1 = fmaxf(fmaxf(red[0][1], red[1][1]), fmaxf(red[2][1], red[3][1]));
  __syncthreads();
  int XS = 1, XE = 0;
  if (U0 <= U1) {
    XS = max(1, (int)floorf(fmaxf(U0 + m, -1e9f)) - 1);
    XE = min(N, (int)ceilf(fminf(U1 + m, 1e9f)) + 1);
  }

  const ACU_GLOBAL float* __restrict__ img = as_global(tk.img);
  const int r1 = min(N - 1, rb * 64 + 63);
  const FplBoxGeom bg = fpl_box_geom(tk, ab * 4 * AL, 4 * AL, rb * 64, r1, lane);
  int c0 = 0, c1 = -1;

  // stage the box B of the chunk [c0, c1] into LDS (all four waves);
  // returns whether this thread staged a value other than +-0
  auto stage = [&](const FplBox& B) {
    unsigned nzb = 0u;  // OR of the staged values' bits without the sign
    const int bx0 = B.x0, by0 = B.y0, bw = B.w, bh = B.h;
    // row staging: lanes cover 128 columns, the wave pair (wave >> 1) picks
    // the row parity, FPL_SB loads in flight per lane; 32-bit byte offsets
    // from the image base and LDS stores at immediate offsets
    const int rp = __builtin_amdgcn_readfirstlane(wave >> 1);
    const int col = lane + ((wave & 1) << 6);
    const int gx = bx0 + col;
    const bool cok = col < bw;
    const ACU_GLOBAL char* ib = (const ACU_GLOBAL char*)img;
    unsigned goff = 4u * (unsigned)((by0 + rp - 1) * N + (gx - 1));  // wraps for rows above the image
    const unsigned gstep = 8u * (unsigned)N;
    float* lrow = box + (rp * PITCH + col);
    if (bx0 >= 1 && by0 >= 1 && bx0 + bw - 1 <= N && by0 + bh - 1 <= N) {
      // the box lies in the image: no per-row tests, whole groups of FPL_SB
      // rows, then one group whose rows past the last repeat the last (the
      // same value to the same address)
      if (cok) {
        const int nrow = (bh - rp + 1) >> 1;  // rows rp, rp + 2, ... < bh
        int g = 0;
        for (; g + FPL_SB <= nrow; g += FPL_SB) {
          float v[FPL_SB];
#pragma unroll
          for (int k = 0; k < FPL_SB; ++k) v[k] = *(const ACU_GLOBAL float*)(ib + (goff + (unsigned)k * gstep));
#pragma unroll
          for (int k = 0; k < FPL_SB; ++k) {
            lrow[2 * k * PITCH] = v[k];
            nzb |= __float_as_uint(v[k]) << 1;
          }
          goff += (unsigned)FPL_SB * gstep;
          lrow += 2 * FPL_SB * PITCH;
        }
        if (g < nrow) {
          const int last = nrow - 1 - g;
          float v[FPL_SB];
#pragma unroll
          for (int k = 0; k < FPL_SB; ++k)
            v[k] = *(const ACU_GLOBAL float*)(ib + (goff + (unsigned)min(k, last) * gstep));
#pragma unroll
          for (int k = 0; k < FPL_SB; ++k) {
            lrow[2 * min(k, last) * PITCH] = v[k];
            nzb |= __float_as_uint(v[k]) << 1;
          }
        }
      }
    } else {
      const bool xin = cok && gx >= 1 && gx <= N;
      for (int ly0 = rp; ly0 < bh; ly0 += 2 * FPL_SB) {
        float v[FPL_SB];
#pragma unroll
        for (int k = 0; k < FPL_SB; ++k) {
          const int gy = by0 + ly0 + 2 * k;
          v[k] = 0.0f;
          if (xin && ly0 + 2 * k < bh && gy >= 1 && gy <= N)
            v[k] = *(const ACU_GLOBAL float*)(ib + (goff + (unsigned)k * gstep));
        }
#pragma unroll
        for (int k = 0; k < FPL_SB; ++k) {
          if (cok && ly0 + 2 * k < bh) lrow[2 * k * PITCH] = v[k];
          nzb |= __float_as_uint(v[k]) << 1;
        }
        goff += (unsigned)FPL_SB * gstep;
        lrow += 2 * FPL_SB * PITCH;
      }
    }
    return nzb != 0u;
  };
  // this wave's samples of the chunk from the staged box B
  auto compute = [&](const FplBox& B) {
    const int bx0 = B.x0, by0 = B.y0, bw = B.w, bh = B.h;
    const bool interior = bx0 >= 2 && by0 >= 2 && bx0 + bw - 1 <= N && by0 + bh - 1 <= N;
    if (interior) {
      // every cell in the box passes the bounds test: fast positions, no
      // branches; an inactive (ray, angle) reads a fixed in-box pixel and
      // its sum is never stored
      float bxl[AL], byl[AL], cj[AL], sj[AL];
#pragma unroll
      for (int j = 0; j < AL; ++j) {
        cj[j] = act[j] ? c[j] : 0.0f;
        sj[j] = act[j] ? s[j] : 0.0f;
        bxl[j] = act[j] ? (m - sy[j]) - (float)bx0 : 1.0f;
        byl[j] = act[j] ? (m + cyv[j]) - (float)by0 : 1.0f;
      }
      // Software-pipelined over two groups of the lane's angles: a group's
      // taps for the next step are read from LDS right after its lerps of
      // this step, so they are in flight while the other group's lerps run
      // (a scheduling barrier keeps the compiler from merging the groups
      // back into one wait).  Each angle still adds its samples in step
      // order.  u = x - m is an integer below 2^24 in magnitude, so
      // stepping it by 1.0f gives exactly (float)x - m without a conversion.
      constexpr int NG = AL < 2 ? AL : 2;  // angle groups
      constexpr int H = AL / NG;           // angles per group
      struct Taps { float fx, fy, v00, v10, v01, v11; };
      Taps tp[NG][H];
      auto fetch = [&](float uu, int g) {
#pragma unroll
        for (int q = 0; q < H; ++q) {
          const int j = g * H + q;
          const float xl = fmaf(cj[j], uu, bxl[j]);
          const float yl = fmaf(sj[j], uu, byl[j]);
          tp[g][q].fx = __builtin_amdgcn_fractf(xl);
          tp[g][q].fy = __builtin_amdgcn_fractf(yl);
          const float* p0 = (const float*)((const char*)box + fpl_cell_offset<PITCH>(xl, yl));
          tp[g][q].v00 = p0[0];
          tp[g][q].v10 = p0[1];
          tp[g][q].v01 = p0[PITCH];
          tp[g][q].v11 = p0[PITCH + 1];
        }
      };
      auto sample = [&](int g) {
#pragma unroll
        for (int q = 0; q < H; ++q) {
          const Taps& t = tp[g][q];
          const float top = fmaf(t.fx, t.v10 - t.v00, t.v00);
          const float bot = fmaf(t.fx, t.v11 - t.v01, t.v01);
          acc[g * H + q] += fmaf(t.fy, bot - top, top);
        }
      };
      float u = (float)c0 - m;
#pragma unroll
      for (int g = 0; g < NG; ++g) fetch(u, g);
      for (int x = c0; x < c1; ++x, u += 1.0f) {
#pragma unroll
        for (int g = 0; g < NG; ++g) {
          sample(g);
          fetch(u + 1.0f, g);
          __builtin_amdgcn_sched_barrier(0);
        }
      }
#pragma unroll
      for (int g = 0; g < NG; ++g) sample(g);  // step c1
    } else {
      // border chunk: the reference's positions and bounds test
      for (int x = c0; x <= c1; ++x) {
        const float u = (float)x - m;
#pragma unroll
        for (int j = 0; j < AL; ++j) {
          if (!act[j]) continue;
          const float xr = (c[j] * u - sy[j]) + m;
          const float yr = (s[j] * u + cyv[j]) + m;
          const float fxr = floorf(xr), fyr = floorf(yr);
          const int x1 = (int)fxr, y1 = (int)fyr;
          if (x1 >= 2 && x1 <= N - 1 && y1 >= 2 && y1 <= N - 1) {
            const float fx = xr - fxr, fy = yr - fyr;
            const float* p0 = box + (__mul24(y1 - by0, PITCH) + (x1 - bx0));
            const float v00 = p0[0], v10 = p0[1], v01 = p0[PITCH], v11 = p0[PITCH + 1];
            const float top = fmaf(fx, v10 - v00, v00);
            const float bot = fmaf(fx, v11 - v01, v01);
            acc[j] += fmaf(fy, bot - top, top);
          }
        }
      }
    }
  };
  // this wave's samples of the chunk, bilinear cells from global memory
  auto gather = [&]() {
    const int P = N + 1;
    const float fN = (float)N;
    for (int x = c0; x <= c1; ++x) {
      const float u = (float)x - m;
#pragma unroll
      for (int j = 0; j < AL; ++j) {
        if (!act[j]) continue;
        const float xr = (c[j] * u - sy[j]) + m;
        const float yr = (s[j] * u + cyv[j]) + m;
        const int cx = (int)__builtin_amdgcn_fmed3f(xr, 0.0f, fN);
        const int cyi = (int)__builtin_amdgcn_fmed3f(yr, 0.0f, fN);
        const float wx = __builtin_amdgcn_fractf(xr), wy = __builtin_amdgcn_fractf(yr);
        const float4 q = global_f4(as_global(tk.cells), (unsigned)(__mul24(cyi, P) + cx));
        const float t1 = fmaf(wx, q.w, q.z);
        const float t2 = fmaf(wx, q.y, q.x);
        acc[j] += fmaf(wy, t1, t2);
      }
    }
  };
  auto fits = [](const FplBox& B) { return B.w > 0 && B.h > 0 && B.w <= PITCH && B.h * PITCH <= FPL_CAP; };

  int kst = 0;  // chunks staged so far
  for (c0 = XS; c0 <= XE; c0 += FPL_U) {
    c1 = min(XE, c0 + FPL_U - 1);
    // bounding box of the chunk's samples (fpl_box: wave-uniform, no LDS)
    const FplBox B = fpl_box(bg, m, c0, c1);
    if (!B.any) continue;  // no active lane in this workgroup (uniform)
    if (fits(B)) {
      // A box of zeros (+-0) adds exactly +0 to every sum (with taps of
      // either sign, fma(fx, v10 - v00, v00) and the lerps after it are +0,
      // and a sum started at +0 never holds -0), so its samples are
      // skipped: TV iterates are zero outside the reconstruction disk until
      // the Laplacian spreads them, a pixel per CG step.  The flag of staged
      // chunk k is red[0][2 + (k & 1)]; thread 0 clears the other one, which
      // nobody reads after the previous chunk's last barrier and nobody sets
      // before this chunk's last barrier.
      float* flag = &red[0][2 + (kst & 1)];
      if (stage(B)) *flag = 1.0f;
      __syncthreads();
      const bool nonzero = *flag != 0.0f;
      if (threadIdx.x == 0) red[0][2 + ((kst + 1) & 1)] = 0.0f;
      if (nonzero) compute(B);
      __syncthreads();
      ++kst;
    } else {
      gather();  // box too large for LDS
    }
  }
#pragma unroll
  for (int j = 0; j < AL; ++j)
    if (act[j]) tk.sino[(size_t)ang[j] * N + ray] = acc[j] / (float)N;
}

template <int AL>
__global__ __launch_bounds__(FPL_THREADS) void fpl_kernel(const FPLTask* __restrict__ tasks, int ntasks,
                                                          int base, int nblocks,
                                                          const int* __restrict__ order) {
  __shared__ float box[FPL_CAP];
  __shared__ float red[4][4];  // per-wave min x, min y, max x, max y (or step range)
  const int lbk = xcd_remap(blockIdx.x, gridDim.x);
  if (lbk >= nblocks) return;
  const int gb = base + (order ? order[lbk] : lbk);
  const FPLTask& tk = tasks[fpl_find_task(tasks, ntasks, gb)];
  const int ab = (gb - tk.block_begin) % tk.nab;
  switch (tk.pclass ? (int)tk.pclass[ab] : FPL_PCLASS_DEFAULT) {
    case 0: fpl_body<AL, FPL_PITCHES[0]>(tk, gb, box, red); break;
    case 1: fpl_body<AL, FPL_PITCHES[1]>(tk, gb, box, red); break;
    case 2: fpl_body<AL, FPL_PITCHES[2]>(tk, gb, box, red); break;
    case 4: fpl_body<AL, FPL_PITCHES[4]>(tk, gb, box, red); break;
    case 5: fpl_body<AL, FPL_PITCHES[5]>(tk, gb, box, red); break;
    case 6: fpl_body<AL, FPL_PITCHES[6]>(tk, gb, box, red); break;
    case 7: fpl_body<AL, FPL_PITCHES[7]>(tk, gb, box, red); break;
    default: fpl_body<AL, FPL_PITCHES[FPL_PCLASS_DEFAULT]>(tk, gb, box, red); break;
  }
}

}  // namespace

int fpl_angles_per_lane(int N, int T) {
  // keep the spread of a 4*al angle block (radius ~0.75 * N/sqrt 2) within
  // what the LDS box leaves after the 64-ray x FPL_U-step footprint;
  // al in {1, 2, 4}
  const double dth = 3.141592653589793 / (double)T;
  const double r = 0.75 * 0.7071 * (double)N;
  const double room = (double)FPL_PITCH - 0.7071 * (64 + FPL_U) - 6.0;
  int al = 1;
  while (al < AMAX && (4.0 * (2 * al) - 1.0) * dth * r <= room) al *= 2;
  return al;
}
// Bank cost of the interior loop's LDS reads for angle block [a0, a1] at
// row pitch P: the two ds_read2_b32 of each of the al instructions of one
// wave (16 neighbouring rays x 4 consecutive angles, the kernel's lane map),
// at two fixed positions of the image.  Each ds_read2_b32 is costed as two
// passes (its first dword, then its second) of each half-wave: per pass, the
// largest number of distinct dwords one of the 32 banks serves (lanes that
// read the same dword share it).  Timed on gfx950 for 180 angles x 8 pitches
// with one angle per instruction, this per-pass cost orders the pitches as
// the measured read time does (correlation 1.00 per angle).
static int pitch_cost(int P, int N, const float* theta_deg, int a0, int a1, int al) {
  static const double pos[2][2] = {{0.31, 0.23}, {-0.52, 0.37}};  // (step, ray) / (N/2) from the centre
  const double pi = 3.141592653589793, h = 0.5 * N, m = 0.5 * (N + 1);
  double sn[16], cs[16];
  for (int k = 0; k < 4 * al; ++k) {
    const double th = pi * ((double)theta_deg[std::min(a1, a0 + k)] - 90.0) / 180.0;
    sn[k] = std::sin(th);
    cs[k] = std::cos(th);
  }
  int cost = 0;
  for (const auto& q : pos) {
    const double u = q[0] * h;
    const int r0 = (int)std::lround(q[1] * h + m) - 8;
    for (int j = 0; j < al; ++j) {
      int addr[64];
      for (int l = 0; l < 64; ++l) {
        const int k = (l >> 4) + 4 * j;
        const double ym = (double)(r0 + (l & 15) + 1) - m;
        // + 4096: positive addresses, the same banks up to a rotation
        const double x = cs[k] * u - sn[k] * ym + m + 4096.0, y = sn[k] * u + cs[k] * ym + m + 4096.0;
        addr[l] = (int)std::floor(y) * P + (int)std::floor(x);
      }
      for (int rd = 0; rd < 2; ++rd)
        for (int e = 0; e < 2; ++e)
          for (int hw = 0; hw < 2; ++hw) {
            int d[32], nd = 0;
            for (int k = 32 * hw; k < 32 * hw + 32; ++k) {
              const int a = addr[k] + rd * P + e;
              bool seen = false;
              for (int t = 0; t < nd; ++t) seen = seen || d[t] == a;
              if (!seen) d[nd++] = a;
            }
            int cnt[32] = {0}, mx = 0;
            for (int t = 0; t < nd; ++t) mx = std::max(mx, ++cnt[d[t] % 32]);
            cost += mx;
          }
    }
  }
  return cost;
}

void fpl_pitch_classes(int N, int T, const float* theta_deg, int al, uint8_t* out) {
  if (al != 1 && al != 2 && al != 4) throw Error("fpl_pitch_classes: al must be 1, 2 or 4");
  const double pi = 3.141592653589793;
  const double r = 0.70711 * ((double)N + 2.0);
  const int ab = 4 * al;
  for (int a0 = 0, blk = 0; a0 < T; a0 += ab, ++blk) {
    const int a1 = std::min(T, a0 + ab) - 1;
    double lo = 1e300, hi = -1e300, ms = 0.0, mc = 0.0;
    for (int a = a0; a <= a1; ++a) {
      const double ct = pi * ((double)theta_deg[a] - 90.0) / 180.0;
      lo = std::min(lo, ct);
      hi = std::max(hi, ct);
      ms = std::max(ms, std::fabs(std::sin(ct)));
      mc = std::max(mc, std::fabs(std::cos(ct)));
    }
    // per-axis box of a chunk (as in fpl_may_overflow, per axis): a fixed
    // footprint plus the drift of a sample across the block's angles, which
    // grows with the distance from the centre (at most `drift`); a pitch
    // tolerates the drift min(P - bx, rows - by) before a chunk overflows
    const double drift = (hi - lo + 1e-6) * r;
    const double bx = 63.0 * ms + (FPL_U - 1) * mc + 8.0, by = 63.0 * mc + (FPL_U - 1) * ms + 8.0;
    auto tolerance = [&](int P) { return std::min((double)P - bx, (double)(FPL_CAP / P) - by); };
    const double base_tol = std::min(tolerance(FPL_PITCHES[FPL_PCLASS_DEFAULT]), drift);
    int best = FPL_PCLASS_DEFAULT, best_cost = pitch_cost(FPL_PITCHES[FPL_PCLASS_DEFAULT], N, theta_deg, a0, a1, al);
    for (int k = 0; k < FPL_NPITCH; ++k) {
      const int P = FPL_PITCHES[k];
      // no more overflowing chunks than the default pitch
      if (k == FPL_PCLASS_DEFAULT || tolerance(P) < base_tol) continue;
      const int cost = pitch_cost(P, N, theta_deg, a0, a1, al);
      if (cost < best_cost) { best = k; best_cost = cost; }
    }
    out[blk] = (uint8_t)best;
  }
}

bool fpl_may_overflow(int N, int T, const float* theta_deg, int al) {
  if (al == FPV_CELL) return fplc_may_overflow(N, T, theta_deg);
  // The box of a chunk spans (per axis) at most the 64-ray x FPL_U-step
  // footprint of one angle, hypot(63, FPL_U - 1), plus the drift of a
  // sample between the first and last angle of its block, span * r with r
  // the largest distance of an in-image sample from the centre, plus the
  // kernel's margins (2 + 4 pixels and the floors).  When that fits the box
  // for every angle block, no chunk takes the global-cells path.
  const double pi = 3.141592653589793;
  const double r = 0.70711 * ((double)N + 2.0);
  const double foot = std::hypot(63.0, (double)(FPL_U - 1));
  const int ab = 4 * al;
  double span = 0.0;
  for (int a0 = 0; a0 < T; a0 += ab) {
    double lo = 1e300, hi = -1e300;
    for (int a = a0; a < std::min(T, a0 + ab); ++a) {
      const double ct = pi * ((double)theta_deg[a] - 90.0) / 180.0;
      lo = std::min(lo, ct);
      hi = std::max(hi, ct);
    }
    span = std::max(span, hi - lo);
  }
  const double extent = foot + (span + 1e-6) * r + 8.0;
  return extent > (double)std::min(FPL_PITCH, FPL_CAP / FPL_PITCH);
}

int fpl_blocks(int N, int T, int al) {
  if (al == FPV_CELL) return fplc_blocks(N, T);
  const int nrb = (N + 63) / 64;
  const int nab = (T + 4 * al - 1) / (4 * al);
  return nrb * nab;
}

void launch_forward_lds(const FPLTask* d_tasks, int ntasks, int base, int blocks, int al,
                        const int* order, hipStream_t s) {
  if (blocks <= 0) return;
  if (al == FPV_CELL) { launch_forward_cell(d_tasks, ntasks, base, blocks, s); return; }
  const dim3 g(xcd_grid(blocks)), t(FPL_THREADS);
  if (al >= 4)
    hipLaunchKernelGGL(fpl_kernel<4>, g, t, 0, s, d_tasks, ntasks, base, blocks, order);
  else if (al == 2)
    hipLaunchKernelGGL(fpl_kernel<2>, g, t, 0, s, d_tasks, ntasks, base, blocks, order);
  else
    hipLaunchKernelGGL(fpl_kernel<1>, g, t, 0, s, d_tasks, ntasks, base, blocks, order);
}

namespace {
// Length of the part of the line {d*n + t*u}, u = (cos a, sin a),
// n = (-sin a, cos a), inside the square [-h, h]^2 (slab intersection).
double square_chord(double d, double a, double h) {
  const double ux = std::cos(a), uy = std::sin(a), px = -d * uy, py = d * ux;
  double lo = -1e300, hi = 1e300;
  for (int k = 0; k < 2; ++k) {
    const double u = k ? uy : ux, p = k ? py : px;
    if (std::fabs(u) < 1e-12) {
      if (std::fabs(p) > h) return 0.0;
      continue;
    }
    const double t0 = (-h - p) / u, t1 = (h - p) / u;
    lo = std::max(lo, std::min(t0, t1));
    hi = std::min(hi, std::max(t0, t1));
  }
  return std::max(0.0, hi - lo);
}
}  // namespace

// A workgroup's time is about proportional to its rays' steps inside the
// image, so a ray block's blocks are weighted by (rays) x (mean chord of its
// middle ray through the valid cells, over 32 angles in [0, pi)).  When the
// natural order already gives the shares about equal work (within 2%: e.g.
// one whole slice per XCD), it is kept (an empty result): an XCD then works
// on one image, and that measured faster (20.04-20.09 against 20.25-20.29 ms
// per 8-slice 2049^2 x 1500 launch inside 20-outer TV).  Otherwise, greedy
// (longest processing time first): ray blocks by decreasing weight, each to
// the share with the least work so far that has room (split over two shares
// when it does not fit), so the shares carry about the same work; a share
// runs its longest workgroups first (the short ones last shorten the time
// its CUs wait for the last one) unless the launch is long (below).
std::vector<int> fpl_dispatch_order(const std::vector<FplGeom>& g, int base, int n) {
  std::vector<int> order((size_t)std::max(n, 0));
  if (n <= 0) return order;
  struct Seg { int first, count; double w; };
  std::vector<Seg> segs;
  for (const FplGeom& t : g) {
    const int nrb = (t.N + 63) / 64;
    const double m = 0.5 * (t.N + 1), h = 0.5 * (t.N - 2);
    for (int rb = 0; rb < nrb; ++rb) {
      const int lo = std::max(base, t.block_begin + rb * t.nab);
      const int hi = std::min(base + n, t.block_begin + (rb + 1) * t.nab);
      if (lo >= hi) continue;
      const int rays = std::min(64, t.N - rb * 64);
      const double d = (rb * 64 + 0.5 * (rays - 1) + 1.0) - m;
      double chord = 0.0;
      for (int k = 0; k < 32; ++k) chord += square_chord(d, (k + 0.5) * (3.141592653589793 / 32), h);
      segs.push_back({lo - base, hi - lo, rays * chord / 32});
    }
  }
  const int C = xcd_grid(n) / XCDS;
  int cap[XCDS];
  for (int k = 0; k < XCDS; ++k) cap[k] = std::max(0, std::min(C, n - k * C));
  // the natural order's shares (xcd_remap alone): kept when balanced
  {
    double nat[XCDS] = {}, tot = 0.0, mx = 0.0;
    int used = 0;
    for (const Seg& sg : segs)
      for (int k = sg.first / C; k < XCDS && k * C < sg.first + sg.count; ++k) {
        const int lo = std::max(sg.first, k * C), hi = std::min(sg.first + sg.count, k * C + cap[k]);
        if (hi > lo) nat[k] += (hi - lo) * sg.w;
      }
    for (int k = 0; k < XCDS; ++k)
      if (cap[k] > 0) { tot += nat[k]; mx = std::max(mx, nat[k]); ++used; }
    if (mx <= 1.02 * tot / used) return {};
  }
  std::stable_sort(segs.begin(), segs.end(), [](const Seg& a, const Seg& b) { return a.w > b.w; });
  std::vector<int> bins[XCDS];
  double load[XCDS] = {};
  for (const Seg& sg : segs) {
    int first = sg.first, left = sg.count;
    while (left > 0) {
      int best = -1;
      for (int k = 0; k < XCDS; ++k)
        if ((int)bins[k].size() < cap[k] && (best < 0 || load[k] < load[best])) best = k;
      if (best < 0) throw Error("fpl_dispatch_order: blocks outside the slices' ranges");
      const int take = std::min(left, cap[best] - (int)bins[best].size());
      for (int i = 0; i < take; ++i) bins[best].push_back(first + i);
      load[best] += take * sg.w;
      first += take;
      left -= take;
    }
  }
  // A launch of many workgroups per slot (the 4-slice 2049^2 launches: 12)
  // runs each share shortest first, a short one longest first: measured on
  // MI355X, shortest first was faster with 12 workgroups per slot (2.68
  // against 2.83 ms per slice on ones-filled images, 10.48 against 10.53 ms
  // per launch inside 50-outer TV) and slower with 3 (1.77 against 1.67 ms).
  if (n >= 6 * FPL_SLOTS)
    for (int k = 0; k < XCDS; ++k) {
      std::vector<std::vector<int>> runs;  // the share's ray-block runs
      for (size_t i = 0; i < bins[k].size(); ++i) {
        if (i == 0 || bins[k][i] != bins[k][i - 1] + 1) runs.emplace_back();
        runs.back().push_back(bins[k][i]);
      }
      bins[k].clear();
      for (size_t r = runs.size(); r-- > 0;) bins[k].insert(bins[k].end(), runs[r].begin(), runs[r].end());
    }
  size_t pos = 0;
  for (int k = 0; k < XCDS; ++k) {
    if ((int)bins[k].size() != cap[k]) throw Error("fpl_dispatch_order: the slices do not cover the launch");
    for (int v : bins[k]) order[pos++] = v;
  }
  return order;
}

}  // namespace acu
