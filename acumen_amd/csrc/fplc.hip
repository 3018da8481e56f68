// fplc.hip -- forward projection (radon, tomography/tv_hpc.F90:129-159) with
// the image staged in LDS as bilinear cells.
//
// Same definition, blocks, chunks and boxes as fpl.hip with 4 angles per
// lane: one workgroup = 64 neighbouring rays x 16 neighbouring angles, the
// sweep in chunks of FPL_U steps, a per-chunk bounding box of the samples
// with two pixels of margin.  What differs is the layout of the box and the
// shape of the workgroup:
//
//  * every LDS element is the 2x2 bilinear cell of its pixel, stored as
//    (v00, v10 - v00, v01, v11 - v01), 16 bytes, so a sample is one
//    ds_read_b128 instead of two ds_read2_b32.  On gfx950 a 16-byte read
//    per lane at the projector's scattered addresses moves 1.2-2.5x the
//    data per clock of the two 8-byte reads (measured).  The differences
//    are formed once per staged pixel, so a sample needs two FMAs for the
//    rows instead of two subtractions and two FMAs, and the result is the
//    same FP32 value as fpl.hip's: top = fma(fx, v10 - v00, v00),
//    bot = fma(fx, v11 - v01, v01), sample = fma(fy, bot - top, top);
//  * a cell box of FPLC_PITCH x FPLC_ROWS cells takes 158 KB, so a
//    workgroup is 16 waves (one angle each) and owns the CU's LDS: the
//    occupancy (16 waves per CU) is that of four 4-wave fpl workgroups.
//
// Staging: the 1024 threads cover 2 column groups of 63 cells (a lane also
// loads the next column for its left neighbour) x 8 row groups of 13 cell
// rows (14 pixel rows); a pixel is loaded once, its right neighbour comes
// from the next lane, its lower neighbour from the lane's next row.
//
// Interior chunks (every cell of the box passes the bounds test) use
// positions by FMA from the ray's start; border chunks use the reference's
// FP32 positions and bounds test.  A chunk whose box exceeds the cell box
// stages raw pixels into the same LDS instead (up to C_RP x C_RROWS, read
// as fpl.hip reads them, same arithmetic); only a box beyond that reads the
// global cells of fp.hip.
#include "fpl_common.hpp"

#include <atomic>

#include <algorithm>
#include <cmath>

namespace acu {
namespace {

constexpr int C_WAVES = 16;
constexpr int C_THREADS = 64 * C_WAVES;
constexpr int C_COLS = 63;                 // cells per column group
constexpr int C_RG = 13;                   // cell rows per row group
constexpr int C_TAB = C_THREADS / 16;     // chunk boxes per table (16 threads per chunk)
constexpr int C_RP = 201;                  // raw-pixel box: row pitch (odd; C_RP + 1 <= 255 dwords)
constexpr int C_RROWS = FPLC_CAP * 4 / C_RP;  // rows of a raw box in the cell box's bytes
static_assert(C_RP + 1 <= 255, "ds_read2 offsets");
static_assert(C_RROWS * C_RP <= FPLC_CAP * 4, "raw box within the cell box's LDS");
constexpr int C_LDS = FPLC_CAP * 16 + C_WAVES * 16 + C_TAB * 16;  // dynamic LDS: box, scratch, table
static_assert(C_LDS <= 163840, "LDS of one CU");
static_assert(2 * C_COLS >= FPLC_PITCH, "two column groups cover a box row");
static_assert((C_WAVES / 2) * C_RG >= FPLC_ROWS, "eight row groups cover the box");

// inclusive min / max scan within each row of 16 lanes (lane 16r + 15
// holds the row's result)
__device__ __forceinline__ float row16_scan_min(float v) {
  v = fminf(v, fpl_dpp<0x111>(v));
  v = fminf(v, fpl_dpp<0x112>(v));
  v = fminf(v, fpl_dpp<0x114>(v));
  v = fminf(v, fpl_dpp<0x118>(v));
  return v;
}
__device__ __forceinline__ float row16_scan_max(float v) {
  v = fmaxf(v, fpl_dpp<0x111>(v));
  v = fmaxf(v, fpl_dpp<0x112>(v));
  v = fmaxf(v, fpl_dpp<0x114>(v));
  v = fmaxf(v, fpl_dpp<0x118>(v));
  return v;
}
// the value of the next lane (lane 63 keeps its own)
__device__ __forceinline__ float lane_next(float v) { return fpl_dpp<0x130>(v); }

typedef float cellv __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) const cellv lds_cellv;
// the cell at index i of the box whose LDS byte address is lbase: one
// shift-add on top of the cell index's multiply-add, one ds_read_b128
__device__ __forceinline__ float4 cell_at(unsigned lbase, unsigned i) {
  const cellv q = *reinterpret_cast<lds_cellv*>((size_t)(lbase + (i << 4)));
  return make_float4(q.x, q.y, q.z, q.w);
}

__device__ __forceinline__ void fplc_body(const FPLTask& tk, int gb, float4* __restrict__ box,
                                          float (*red)[4], int4* __restrict__ btab) {
#pragma clang fp contract(off)
  const int N = tk.N, T = tk.T;
  const float m = tk.mid;
  const int local = gb - tk.block_begin;
  const int rb = local / tk.nab, ab = local - rb * tk.nab;  // ray-block-major
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ray = rb * 64 + lane;
  const float ym = (float)(min(ray, N - 1) + 1) - m;
  const int ang = ab * C_WAVES + wave;
  const bool act = ang < T && ray < N;
  const float2 scv = const_f2(tk.sc, min(ang, T - 1));
  const float s = scv.x, c = scv.y;
  const float sy = s * ym, cyv = c * ym;  // one rounding each (tv_hpc.F90:143-144)
  float acc = 0.0f;

  // workgroup step window (1-based steps, one step of margin)
  float lu0 = 1e30f, lu1 = -1e30f;
  if (act) {
    float ua, ub, va, vb;
    fpl_urange(m - sy, c, 1.99f, (float)N + 0.01f, ua, ub);
    fpl_urange(m + cyv, s, 1.99f, (float)N + 0.01f, va, vb);
    const float u0 = fmaxf(ua, va), u1 = fminf(ub, vb);
    if (u0 <= u1) { lu0 = u0; lu1 = u1; }
  }
  lu0 = fpl_wmin(lu0);
  lu1 = fpl_wmax(lu1);
  if (lane == 0) { red[wave][0] = lu0; red[wave][1] = lu1; }
  __syncthreads();
  float U0 = red[0][0], U1 = red[0][1];
#pragma unroll
  for (int w = 1; w < C_WAVES; ++w) { U0 = fminf(U0, red[w][0]); U1 = fmaxf(U1, red[w][1]); }
  __syncthreads();
  int XS = 1, XE = 0;
  if (U0 <= U1) {
    XS = max(1, (int)floorf(fmaxf(U0 + m, -1e9f)) - 1);
    XE = min(N, (int)ceilf(fminf(U1 + m, 1e9f)) + 1);
  }

  const ACU_GLOBAL float* __restrict__ img = as_global(tk.img);
  // staging coordinates of this thread: column group, lane column, row group
  const int scol = (wave & 1) * C_COLS + lane;
  const int srow0 = (wave >> 1) * C_RG;

  // Boxes of the chunks, C_TAB at a time, in a table in LDS: thread t takes
  // chunk kbase + t / 16 and angle t % 16 of the block, the four corners of
  // its samples (first / last ray x first / last step), and a DPP scan over
  // each 16-lane row gives the chunk's min / max at the row's last lane,
  // which writes the entry -- the box that reducing over every (ray, angle,
  // step) gives, because positions are monotone in the ray offset and the
  // step (see fpl_box).  1-based pixel origin with two pixels of margin,
  // size in cells; w = 0 when the chunk holds no active (ray, angle).
  struct Box { int x0, y0, w, h; bool fits, raw; };  // fits: cell box; raw: raw-pixel box
  const int ta = ab * C_WAVES + (threadIdx.x & 15);
  const bool tact = ta < T;
  const float2 tsc = const_f2(tk.sc, min(ta, T - 1));
  const int r0 = rb * 64, r1 = min(N - 1, rb * 64 + 63);
  const float tym0 = (float)(r0 + 1) - m, tym1 = (float)(r1 + 1) - m;
  const float tpx0 = m - tsc.x * tym0, tpx1 = m - tsc.x * tym1;  // m - fl(s*ym)
  const float tpy0 = m + tsc.y * tym0, tpy1 = m + tsc.y * tym1;  // m + fl(c*ym)
  const int nch = XE >= XS ? (XE - XS) / FPL_U + 1 : 0;
  auto fill_boxes = [&](int kbase) {
    const int kc = kbase + (int)(threadIdx.x >> 4);
    const int k0 = XS + kc * FPL_U, k1 = min(XE, k0 + FPL_U - 1);
    float mnx = 1e30f, mny = 1e30f, mxx = -1e30f, mxy = -1e30f;
    if (tact && kc < nch) {
      const float ua = (float)k0 - m, ub = (float)k1 - m;
      const float x00 = fmaf(tsc.y, ua, tpx0), x01 = fmaf(tsc.y, ub, tpx0);
      const float x10 = fmaf(tsc.y, ua, tpx1), x11 = fmaf(tsc.y, ub, tpx1);
      const float y00 = fmaf(tsc.x, ua, tpy0), y01 = fmaf(tsc.x, ub, tpy0);
      const float y10 = fmaf(tsc.x, ua, tpy1), y11 = fmaf(tsc.x, ub, tpy1);
      mnx = fminf(fminf(x00, x01), fminf(x10, x11)); mxx = fmaxf(fmaxf(x00, x01), fmaxf(x10, x11));
      mny = fminf(fminf(y00, y01), fminf(y10, y11)); mxy = fmaxf(fmaxf(y00, y01), fmaxf(y10, y11));
    }
    mnx = row16_scan_min(mnx); mny = row16_scan_min(mny);
    mxx = row16_scan_max(mxx); mxy = row16_scan_max(mxy);
    if ((threadIdx.x & 15) == 15) {
      int4 e = make_int4(0, 0, 0, 0);
      if (mnx <= mxx) {
        e.x = (int)floorf(fmaxf(mnx, -1e6f)) - 2;
        e.y = (int)floorf(fmaxf(mny, -1e6f)) - 2;
        e.z = (int)floorf(fminf(mxx, 1e6f)) - e.x + 4;
        e.w = (int)floorf(fminf(mxy, 1e6f)) - e.y + 4;
      }
      btab[threadIdx.x >> 4] = e;
    }
  };
  int kbase = 0;
  auto box_at = [&](int kc) {
    const int4 e = btab[kc - kbase];
    Box bx{e.x, e.y, e.z, e.w, false, false};
    bx.fits = bx.w > 0 && bx.h > 0 && bx.w <= FPLC_PITCH && bx.h <= FPLC_ROWS;
    bx.raw = !bx.fits && bx.w > 0 && bx.h > 0 && bx.w <= C_RP && bx.h <= C_RROWS;
    return bx;
  };
  // This thread's pixels of a box: column scol (and its right neighbour
  // through the next lane), rows srow0 .. srow0 + C_RG.  Rows and columns
  // are clamped into the image and loaded without conditions: a pixel
  // outside the image only enters cells that no sample reads (the bounds
  // test keeps taps in [2, N]; interior boxes lie inside the image).
  float v[C_RG + 1];
  auto load_pixels = [&](const Box& bx) {
    if (!bx.fits) return;  // uniform
    const int gx = min(max(bx.x0 + scol, 1), N) - 1;
#pragma unroll
    for (int k = 0; k <= C_RG; ++k) {
      const int gy = min(max(bx.y0 + srow0 + k, 1), N) - 1;  // wave-uniform
      v[k] = img[(size_t)gy * N + gx];
    }
  };

  // Software pipeline: the pixels of chunk k+1 are loaded into registers
  // while chunk k is sampled from LDS, and written as cells after the
  // barrier that ends chunk k.
  fill_boxes(0);
  __syncthreads();
  Box cur = nch > 0 ? box_at(0) : Box{0, 0, 0, 0, false, false};
  load_pixels(cur);
  for (int kc = 0; kc < nch; ++kc) {
    const int c0 = XS + kc * FPL_U;
    const int c1 = min(XE, c0 + FPL_U - 1);
    // (1) this chunk's cells from the registers (every wave passed the
    // barrier that ended the previous chunk)
    if (cur.fits) {
      float nb[C_RG + 1];
#pragma unroll
      for (int k = 0; k <= C_RG; ++k) nb[k] = lane_next(v[k]);  // the pixel to the right
      const bool cw = lane < C_COLS && scol < cur.w;
#pragma unroll
      for (int k = 0; k < C_RG; ++k)
        if (cw && srow0 + k < cur.h)
          box[(srow0 + k) * FPLC_PITCH + scol] = make_float4(v[k], nb[k] - v[k], v[k + 1], nb[k + 1] - v[k + 1]);
    } else if (cur.raw) {
      // a box too large for the cells: raw pixels, row pitch C_RP, staged
      // here without the prefetch (4 column groups of 64 x rows by wave
      // quarter, 8 loads in flight; clamped loads as above)
      float* rbox = reinterpret_cast<float*>(box);
      const int col = (wave & 3) * 64 + lane;
      const int gx = min(max(cur.x0 + col, 1), N) - 1;
      const bool cok = col < cur.w;
      for (int r0w = wave >> 2; r0w < cur.h; r0w += 4 * 8) {
        float t[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const int gy = min(max(cur.y0 + r0w + 4 * k, 1), N) - 1;  // wave-uniform
          t[k] = img[(size_t)gy * N + gx];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k)
          if (cok && r0w + 4 * k < cur.h) rbox[(r0w + 4 * k) * C_RP + col] = t[k];
      }
    }
    if (kc + 1 == kbase + C_TAB && kc + 1 < nch) {  // uniform: the next table of boxes
      kbase += C_TAB;
      fill_boxes(kbase);
    }
    __syncthreads();
    // (2) the next chunk's box, and its pixels in flight during (3)
    const Box nxt = kc + 1 < nch ? box_at(kc + 1) : Box{0, 0, 0, 0, false, false};
    load_pixels(nxt);
    // (3) sample this chunk
    const int bx0 = cur.x0, by0 = cur.y0, bw = cur.w, bh = cur.h;
    if (cur.fits) {
      const bool interior = bx0 >= 2 && by0 >= 2 && bx0 + bw - 1 <= N && by0 + bh - 1 <= N;
      if (interior) {
        // every cell in the box passes the bounds test: fast positions, no
        // branches; an inactive (ray, angle) reads a fixed in-box cell and
        // its sum is never stored
        const float cj = act ? c : 0.0f, sj = act ? s : 0.0f;
        const float bxl = act ? (m - sy) - (float)bx0 : 1.0f;
        const float byl = act ? (m + cyv) - (float)by0 : 1.0f;
        const unsigned lbase = (unsigned)(uintptr_t)box;  // LDS byte address of the box (uniform)
        float u = (float)c0 - m;  // integer-valued, so u += 1 is exact
#pragma unroll 4
        for (int x = c0; x <= c1; ++x, u += 1.0f) {
          const float xl = fmaf(cj, u, bxl);
          const float yl = fmaf(sj, u, byl);
          const int ix = (int)xl, iy = (int)yl;
          const float fx = __builtin_amdgcn_fractf(xl), fy = __builtin_amdgcn_fractf(yl);
          const float4 q = cell_at(lbase, __umul24((unsigned)iy, FPLC_PITCH) + (unsigned)ix);
          const float top = fmaf(fx, q.y, q.x);
          const float bot = fmaf(fx, q.w, q.z);
          acc += fmaf(fy, bot - top, top);
        }
      } else if (act) {
        // border chunk: the reference's positions and bounds test
        for (int x = c0; x <= c1; ++x) {
          const float u = (float)x - m;
          const float xr = (c * u - sy) + m;
          const float yr = (s * u + cyv) + m;
          const float fxr = floorf(xr), fyr = floorf(yr);
          const int x1 = (int)fxr, y1 = (int)fyr;
          if (x1 >= 2 && x1 <= N - 1 && y1 >= 2 && y1 <= N - 1) {
            const float fx = xr - fxr, fy = yr - fyr;
            const float4 q = box[__mul24(y1 - by0, FPLC_PITCH) + (x1 - bx0)];
            const float top = fmaf(fx, q.y, q.x);
            const float bot = fmaf(fx, q.w, q.z);
            acc += fmaf(fy, bot - top, top);
          }
        }
      }
    } else if (cur.raw) {
      const float* rbox = reinterpret_cast<const float*>(box);
      const bool interior = bx0 >= 2 && by0 >= 2 && bx0 + bw - 1 <= N && by0 + bh - 1 <= N;
      if (interior) {
        // fpl.hip's fast path on raw pixels (same arithmetic as the cells)
        const float cj = act ? c : 0.0f, sj = act ? s : 0.0f;
        const float bxl = act ? (m - sy) - (float)bx0 : 1.0f;
        const float byl = act ? (m + cyv) - (float)by0 : 1.0f;
        float u = (float)c0 - m;
#pragma unroll 4
        for (int x = c0; x <= c1; ++x, u += 1.0f) {
          const float xl = fmaf(cj, u, bxl);
          const float yl = fmaf(sj, u, byl);
          const int ix = (int)xl, iy = (int)yl;
          const float fx = __builtin_amdgcn_fractf(xl), fy = __builtin_amdgcn_fractf(yl);
          const float* p0 = rbox + (__mul24(iy, C_RP) + ix);
          const float v00 = p0[0], v10 = p0[1], v01 = p0[C_RP], v11 = p0[C_RP + 1];
          const float top = fmaf(fx, v10 - v00, v00);
          const float bot = fmaf(fx, v11 - v01, v01);
          acc += fmaf(fy, bot - top, top);
        }
      } else if (act) {
        for (int x = c0; x <= c1; ++x) {
          const float u = (float)x - m;
          const float xr = (c * u - sy) + m;
          const float yr = (s * u + cyv) + m;
          const float fxr = floorf(xr), fyr = floorf(yr);
          const int x1 = (int)fxr, y1 = (int)fyr;
          if (x1 >= 2 && x1 <= N - 1 && y1 >= 2 && y1 <= N - 1) {
            const float fx = xr - fxr, fy = yr - fyr;
            const float* p0 = rbox + (__mul24(y1 - by0, C_RP) + (x1 - bx0));
            const float v00 = p0[0], v10 = p0[1], v01 = p0[C_RP], v11 = p0[C_RP + 1];
            const float top = fmaf(fx, v10 - v00, v00);
            const float bot = fmaf(fx, v11 - v01, v01);
            acc += fmaf(fy, bot - top, top);
          }
        }
      }
    } else if (act) {
      // box too large for LDS: bilinear cells from global memory (as fpl.hip)
      const int P = N + 1;
      const float fN = (float)N;
      for (int x = c0; x <= c1; ++x) {
        const float u = (float)x - m;
        const float xr = (c * u - sy) + m;
        const float yr = (s * u + cyv) + m;
        const int cx = (int)__builtin_amdgcn_fmed3f(xr, 0.0f, fN);
        const int cyi = (int)__builtin_amdgcn_fmed3f(yr, 0.0f, fN);
        const float wx = __builtin_amdgcn_fractf(xr), wy = __builtin_amdgcn_fractf(yr);
        const float4 q = global_f4(as_global(tk.cells), (unsigned)(__mul24(cyi, P) + cx));
        const float t1 = fmaf(wx, q.w, q.z);
        const float t2 = fmaf(wx, q.y, q.x);
        acc += fmaf(wy, t1, t2);
      }
    }
    __syncthreads();  // every wave is done with this chunk's cells
    cur = nxt;
  }
  if (act) tk.sino[(size_t)ang * N + ray] = acc / (float)N;
}

__global__ __launch_bounds__(C_THREADS) void fplc_kernel(const FPLTask* __restrict__ tasks, int ntasks, int base,
                                                          int nblocks) {
  // dynamic LDS: the cell box at address 0 (so a cell's address is one
  // multiply-add), then per-wave min / max scratch, then the table of boxes
  extern __shared__ float4 cbox[];
  float (*red)[4] = reinterpret_cast<float (*)[4]>(cbox + FPLC_CAP);
  int4* btab = reinterpret_cast<int4*>(cbox + FPLC_CAP + C_WAVES);
  const int lbk = xcd_remap(blockIdx.x, gridDim.x);
  if (lbk >= nblocks) return;
  const int gb = base + lbk;
  fplc_body(tasks[fpl_find_task(tasks, ntasks, gb)], gb, cbox, red, btab);
}

}  // namespace

bool fplc_may_overflow(int N, int T, const float* theta_deg) {
  // as fpl_may_overflow for 16-angle blocks; a chunk reads the global cells
  // only when its box exceeds the raw-pixel box as well
  const double pi = 3.141592653589793;
  const double r = 0.70711 * ((double)N + 2.0);
  const double foot = std::hypot(63.0, (double)(FPL_U - 1));
  double span = 0.0;
  for (int a0 = 0; a0 < T; a0 += C_WAVES) {
    double lo = 1e300, hi = -1e300;
    for (int a = a0; a < std::min(T, a0 + C_WAVES); ++a) {
      const double ct = pi * ((double)theta_deg[a] - 90.0) / 180.0;
      lo = std::min(lo, ct);
      hi = std::max(hi, ct);
    }
    span = std::max(span, hi - lo);
  }
  const double extent = foot + (span + 1e-6) * r + 8.0;
  return extent > (double)std::min(C_RP, C_RROWS);  // beyond the raw-pixel box: global cells
}

int fplc_blocks(int N, int T) { return ((N + 63) / 64) * ((T + C_WAVES - 1) / C_WAVES); }

void launch_forward_cell(const FPLTask* d_tasks, int ntasks, int base, int blocks, hipStream_t s) {
  if (blocks <= 0) return;
  int dev = 0;
  ACU_HIP_CHECK(hipGetDevice(&dev));
  // the LDS size above 64 KB needs the attribute once per device; setting it
  // is idempotent, so threads that race here at most set it twice
  static std::atomic<bool> attr_set[64];
  if (dev >= 64 || !attr_set[dev].load(std::memory_order_acquire)) {
    ACU_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&fplc_kernel),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, C_LDS));
    if (dev < 64) attr_set[dev].store(true, std::memory_order_release);
  }
  hipLaunchKernelGGL(fplc_kernel, dim3(xcd_grid(blocks)), dim3(C_THREADS), C_LDS, s, d_tasks, ntasks, base, blocks);
}

}  // namespace acu
