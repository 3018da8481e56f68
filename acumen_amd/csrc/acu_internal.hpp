// acu_internal.hpp -- shared declarations of the gfx950 reconstruction library.
//
// Layout in HBM (per slice, N = image side = detector bins, T = angles):
//   image arrays   N*N fp32, x fastest (the reference's img(x,y), column-major)
//   sinograms      N*T fp32, detector fastest, one row per angle
//   cells          (N+1)*(N+1) float4, bilinear cells of an image (see fp.hip)
//   angle tables   T float2 (sin, cos) of the reference's ctheta, computed on
//                  the host with the reference's FP32 expression
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

// Address-space views for device code: uniform tables read through the
// constant space become scalar (SMEM) loads, data pointers read through the
// global space become global_load with an SGPR base.
#define ACU_GLOBAL __attribute__((address_space(1)))
#define ACU_CONST __attribute__((address_space(4)))

namespace acu {

template <class T>
__device__ __forceinline__ const ACU_CONST T* as_const(const T* p) {
  return (const ACU_CONST T*)p;
}
template <class T>
__device__ __forceinline__ const ACU_GLOBAL T* as_global(const T* p) {
  return (const ACU_GLOBAL T*)p;
}
// float2 element i of a table read through the constant space (HIP's vector
// types cannot be copied out of a non-generic address space directly)
__device__ __forceinline__ float2 const_f2(const float2* tab, size_t i) {
  const ACU_CONST float* p = (const ACU_CONST float*)tab + 2 * i;
  return make_float2(p[0], p[1]);
}
__device__ __forceinline__ float4 global_f4(const ACU_GLOBAL float4* tab, unsigned i) {
  const ACU_GLOBAL float4* p = tab + i;
  typedef float v4 __attribute__((ext_vector_type(4)));
  const v4 v = *(const ACU_GLOBAL v4*)p;
  return make_float4(v.x, v.y, v.z, v.w);
}

// The dispatcher deals workgroups round-robin over the 8 XCDs of MI355X,
// each with its own L2.  xcd_remap turns the hardware block index of a grid
// of xcd_grid(n) blocks into a logical index in [0, xcd_grid(n)) such that
// the blocks one XCD runs at the same time are consecutive logical blocks;
// indices >= n are padding and exit at once.
constexpr int XCDS = 8;
__host__ __device__ inline int xcd_grid(int n) { return (n + XCDS - 1) / XCDS * XCDS; }
__device__ __forceinline__ int xcd_remap(int b, int grid) {
  return (b % XCDS) * (grid / XCDS) + b / XCDS;
}

// ---- backprojection tiling (bp.hip) ----
// A tile is 64 (x) x ty (y) pixels, ty = 16, 32 or 64 (a per-batch choice),
// one workgroup of 4 waves (8 at ty = 64).  Tile activity is kept per
// 64 x 16 sub-tile, so every tile height reads the same table.
constexpr int BP_TX = 64;        // pixels along x per tile (one per lane)
constexpr int BP_SUB = 16;       // rows of one activity sub-tile
constexpr int BP_TY_DEFAULT = 32;
constexpr int BP_AB = 32;        // angles staged per LDS block
inline bool bp_tile_ok(int ty) { return ty == 16 || ty == 32 || ty == 64; }

// ---- forward projection (fp.hip) ----
constexpr int FP_THREADS = 256;  // rays per workgroup (one angle per workgroup)
constexpr int FP_CHUNK_DEFAULT = 1 << 20;  // steps per sweep chunk (default: the whole ray)

// ---- forward projection with LDS-staged boxes (fpl.hip) ----
#ifndef ACUMEN_FPL_U
#define ACUMEN_FPL_U 32
#endif
#ifndef ACUMEN_FPL_PITCH
#define ACUMEN_FPL_PITCH 101
#endif
#ifndef ACUMEN_FPL_ROWS
#define ACUMEN_FPL_ROWS 101
#endif
constexpr int FPL_U = ACUMEN_FPL_U;          // steps per staged chunk
constexpr int FPL_PITCH = ACUMEN_FPL_PITCH;  // LDS row pitch of the box (odd; pitch+1 <= 255 dwords)
constexpr int FPL_CAP = ACUMEN_FPL_PITCH * ACUMEN_FPL_ROWS;  // floats of the LDS box
static_assert(FPL_PITCH + 1 <= 255, "ds_read2 offsets");
// Row pitches an angle block may use for its boxes (the capacity FPL_CAP is
// fixed; rows = FPL_CAP / pitch); class FPL_PCLASS_DEFAULT is FPL_PITCH.
constexpr int FPL_NPITCH = 8;
#ifndef ACUMEN_FPL_PITCH_CLASSES
#define ACUMEN_FPL_PITCH_CLASSES 93, 97, 99, 101, 103, 105, 107, 109
#endif
constexpr int FPL_PITCHES[FPL_NPITCH] = {ACUMEN_FPL_PITCH_CLASSES};
constexpr int FPL_PCLASS_DEFAULT = 3;
// Cell-layout LDS projector (fplc.hip): the box is FPLC_PITCH x FPLC_ROWS
// 16-byte bilinear cells (161,600 bytes of the CU's 160 KiB LDS).
constexpr int FPLC_PITCH = 101;
constexpr int FPLC_ROWS = 100;
constexpr int FPLC_CAP = FPLC_PITCH * FPLC_ROWS;
// Projector variants (per slice): 0 = global gathers (fp.hip), 1/2/4 = LDS
// boxes with that many angles per lane (fpl.hip), FPV_CELL = cell boxes,
// one angle per wave, 16 waves (fplc.hip).
constexpr int FPV_CELL = 8;
static_assert(FPL_PITCHES[FPL_PCLASS_DEFAULT] == FPL_PITCH, "default pitch class");

// ---- ramp filter GEMM (filter.hip) ----
constexpr int RF_BM = 64;   // detector rows per tile
constexpr int RF_BN = 64;   // angles per tile
constexpr int RF_BK = 32;   // reduction step
constexpr int RF_THREADS = 256;

// ---- elementwise / reductions (tv.hip) ----
constexpr int EW_THREADS = 256;
constexpr int EW_ITEMS = 4;      // elements per thread per block pass

// One slice of a batched operator launch.
struct BPTask {
  int N, T;
  float mid;             // (N+1)/2 with integer division
  int tile_begin;        // first global tile index of this slice
  int ntx, nty;          // tiles along x and y (at the launch's tile height)
  int nsub;              // 16-row sub-tiles along y, ceil(N / 16)
  const float* sino;     // filtered sinogram (N*T)
  const float* cmask;    // N*N
  const float2* sc;      // per-angle (sin, cos) for backprojection
  const uint8_t* tile_active;  // nsub*ntx: 1 if any cmask > 0 in the 64 x 16 sub-tile
  const float* rowterm;  // T*N (+ padding): cos(c_t) * (y - midpoint) in FP32, angle-major
  float* img;            // N*N out
};

struct FPTask {
  int N, T;
  float mid;
  int block_begin;       // first global block index of this slice
  int nrb;               // ray blocks per angle = ceil(N / FP_THREADS)
  int nchunk;            // sweep chunks per ray = ceil(N / clen); block = (chunk*nrb + rb)*T + angle
  int clen;              // steps per chunk
  int sum_begin;         // first global block index of this slice in the chunk-sum pass
  float* partial;        // nchunk*T*N per-chunk partial ray sums
  const float4* cells;   // (N+1)^2 cells, x-major
  const float4* cellsT;  // (N+1)^2 cells, y-major (transposed)
  const float2* sc;      // per-angle (sin, cos) for projection
  float* sino;           // N*T out
};

struct FPLTask {
  int N, T;
  float mid;
  int block_begin;       // first global block index of this slice
  int al;                // angles per lane (1, 2 or 4), or FPV_CELL; the same for every slice of a launch
  int nab;               // angle blocks = ceil(T / fpl_angle_block(al)); block = rb * nab + ab
  const float* img;      // N*N image (raw pixels, staged in LDS)
  const float4* cells;   // x-major cells of the same image (fallback for oversized boxes)
  const float2* sc;      // per-angle (sin, cos) for projection
  float* sino;           // N*T out
  const uint8_t* pclass; // nab pitch classes (fpl_pitch_classes), or null: FPL_PITCH
};

struct RFTask {
  int N, T;
  int tile_begin;
  int tm, tn;            // tiles along detector rows and angles
  const float* h;        // 2N+1 taps
  const float* sino;     // N*T in
  float* out;            // N*T out
};

struct CellTask {
  int N;
  int block_begin;
  const float* img;      // N*N in (optionally updated in place, see tv.hip)
  float4* cells;
  float4* cellsT;
};

// Launchers (stream-ordered, no host synchronisation).  A launch covers the
// global blocks [base, base+blocks) of a task array: base = 0 and all blocks
// for a whole batch, or one slice's own range.
// arithmetic of the ramp filter (filter.hip); RAMP_FFT is fft.hip
enum RampVariant { RAMP_FP64 = 0, RAMP_FP32 = 1, RAMP_BF16X3 = 2, RAMP_BF16X6 = 3, RAMP_FFT = 4 };
void launch_ramp(const RFTask* d_tasks, int ntasks, int base, int tiles, int variant, hipStream_t s);

// ---- ramp filter by FFT convolution (fft.hip) ----
constexpr int FFT_MAX_L = 8192;  // transform length limit (LDS: 8 B per point)
struct FFTTask {
  int N, T, L, lg;     // L = 2^lg = fft_len(N)
  int block_begin;     // one block per pair of angle rows
  int alias;           // L == 2N-2: outputs 0 and N-1 get the correction
  float g_edge;        // tap at distance N-1
  const float* G;      // L: spectrum of the wrapped kernel, times 1/L
  const float2* tw;    // L/2: exp(-2 pi i k / L)
  const float* sino;   // (T, N) in
  float* out;          // (T, N) out
};
int fft_len(int N);
bool fft_supported(int N);
void fft_tables(int N, const float* taps, float* G, float2* tw);
void launch_ramp_fft(const FFTTask* d_tasks, int ntasks, int base, int blocks, int maxL, hipStream_t s);
// refpos: positions in the reference's FP32 arithmetic (default); false:
// tile-relative positions (closer to FP64, two VALU fewer per update)
// XCD shares of a backprojection launch over `tiles` tiles: XCD k runs the
// launch's tiles [b[k], b[k+1]) in order (b[XCDS] = tiles), at most C each.
struct BPShares { int b[XCDS + 1]; int C; };
// Shares of about equal work for tiles of estimated work w (launch order):
// contiguous, so an XCD's L2 serves neighbouring tiles of one slice.
BPShares bp_xcd_shares(const std::vector<float>& w);
void launch_backproject(const BPTask* d_tasks, int ntasks, int base, int tiles, int ty, bool refpos,
                        const BPShares& sh, hipStream_t s);
void launch_cells(const CellTask* d_tasks, int ntasks, int base, int blocks, hipStream_t s);
// rowterm[t*N + y] = cos(c_t) * ((y+1) - mid), one FP32 product each (the
// backprojector's row term of tv_hpc.F90:114), for t < T, y < N
constexpr int BP_ROWTERM_PAD = 64;  // floats after the table (the last tile's rows past N)
void launch_bp_rowterm(const float2* sc, int N, int T, float mid, float* rowterm, hipStream_t s);
void launch_forward(const FPTask* d_tasks, int ntasks, int base, int blocks, int sum_base,
                    int sum_blocks, hipStream_t s);
// al = 1, 2, 4: fpl.hip; al = FPV_CELL: fplc.hip.  order (al = 1, 2, 4;
// null: block i is block base + i): the launch's i-th logical block is block
// base + order[i], a permutation of [0, blocks) from fpl_dispatch_order.
void launch_forward_lds(const FPLTask* d_tasks, int ntasks, int base, int blocks, int al,
                        const int* order, hipStream_t s);
void launch_forward_cell(const FPLTask* d_tasks, int ntasks, int base, int blocks, hipStream_t s);

// Host-side helpers.
int bp_tiles_x(int N);
int bp_tiles_y(int N, int ty);
int rf_tiles_m(int N);
int rf_tiles_n(int T);
int fp_ray_blocks(int N);
int fp_chunks(int N, int clen);
int fp_sum_blocks(int N, int T);
int fpl_angles_per_lane(int N, int T);
int fpl_blocks(int N, int T, int al);
// One slice of an LDS projector launch: side N, nab angle blocks, its first
// block (blocks rb * nab + ab from there, rb = ray block of 64 rays).
struct FplGeom { int N, nab, block_begin; };
// Dispatch order of a launch over blocks [base, base + n) of the slices g:
// empty when the natural order gives each XCD's share of the grid (the
// logical blocks [k*C, (k+1)*C), C = xcd_grid(n) / XCDS, that xcd_remap
// hands XCD k) about the same estimated work, else a permutation of [0, n)
// that does, longest workgroups first (shortest first in launches of 6 or
// more workgroups per slot).  A ray block's angle blocks stay consecutive.
std::vector<int> fpl_dispatch_order(const std::vector<FplGeom>& g, int base, int n);
int fplc_blocks(int N, int T);
// angles per workgroup block of an LDS projector variant
inline int fpl_angle_block(int al) { return al == FPV_CELL ? 16 : 4 * al; }
// true when some chunk of the LDS projector can outgrow its box and read
// the global bilinear cells (which then have to be built first)
bool fpl_may_overflow(int N, int T, const float* theta_deg, int al);
bool fplc_may_overflow(int N, int T, const float* theta_deg);
// Pitch class of every block of 4*al angles: the pitch with the fewest LDS
// bank conflicts for the block's angles under the kernel's lane map, among
// those whose rows hold the block's box (out: ceil(T / (4*al)) bytes)
void fpl_pitch_classes(int N, int T, const float* theta_deg, int al, uint8_t* out);
int cell_blocks(int N);

}  // namespace acu

#define ACU_HIP_CHECK(expr)                                                  \
  do {                                                                       \
    hipError_t _e = (expr);                                                  \
    if (_e != hipSuccess) {                                                  \
      throw acu::Error(std::string(#expr) + ": " + hipGetErrorString(_e));   \
    }                                                                        \
  } while (0)

namespace acu {
struct Error {
  std::string msg;
  explicit Error(std::string m) : msg(std::move(m)) {}
};
}  // namespace acu

// Context internals for entry points defined outside api.hip.
struct acu_ctx;
namespace acu {
int ctx_device(const acu_ctx* c);
hipStream_t ctx_stream(const acu_ctx* c);
// Record msg as the calling thread's acu_last_error text; returns code.
int record_error(int code, const std::string& msg);
// Device address of the no-alloc entry point's static scratch on `device`
// (noalloc.hip), looked up once per device with the device current.
void noalloc_setup(int device);
}  // namespace acu
