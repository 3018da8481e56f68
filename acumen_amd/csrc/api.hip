// api.hip -- host runtime and C ABI (include/acumen.h) of the gfx950
// reconstruction library: plans/batches (the reference's module state),
// operator launches, and the split-Bregman TV orchestration.
#include "acu_tv.hpp"
#include "acu_grad.hpp"
#include "acu_prep.hpp"
#include "acumen.h"

#include <chrono>
#include <map>
#include <mutex>
#include <thread>
#include <sstream>
#include <tuple>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

namespace {
thread_local std::string g_err = "no error";
int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
}  // namespace

struct acu_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool timing = false;
  struct Pending { int kind; hipEvent_t a, b; };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  // ramp, backprojection, cells, forward, adjoint, preprocessing, ingest
  static constexpr int KINDS = 7;
  double ms[KINDS] = {0, 0, 0, 0, 0, 0, 0};
  long long count[KINDS] = {0, 0, 0, 0, 0, 0, 0};

  hipEvent_t ev() {
    if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
    hipEvent_t e;
    ACU_HIP_CHECK(hipEventCreate(&e));
    return e;
  }
  // Fold finished timing events into the totals (call after a stream sync).
  void drain() {
    for (auto& p : pending) {
      float t = 0.f;
      ACU_HIP_CHECK(hipEventElapsedTime(&t, p.a, p.b));
      ms[p.kind] += t;
      count[p.kind] += 1;
      pool.push_back(p.a);
      pool.push_back(p.b);
    }
    pending.clear();
  }
  void sync() {
    ACU_HIP_CHECK(hipStreamSynchronize(stream));
    drain();
  }
  // Two pinned host buffers for calls that take pageable host arrays: a host
  // copy into one overlaps the DMA of the other.  Grown on demand (after a
  // stream sync, since a DMA may still use the old ones) and held until the
  // context is destroyed.
  float* pin[2] = {nullptr, nullptr};
  size_t pin_bytes = 0;
  hipEvent_t pin_ev[2] = {nullptr, nullptr};  // last DMA through each buffer
  void pinned(size_t bytes) {
    if (!pin_ev[0])
      for (auto& e : pin_ev) ACU_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    if (bytes <= pin_bytes) return;
    sync();
    for (auto& b : pin) {
      if (b) ACU_HIP_CHECK(hipHostFree(b));
      b = nullptr;
    }
    pin_bytes = 0;
    for (auto& b : pin) ACU_HIP_CHECK(hipHostMalloc((void**)&b, bytes, hipHostMallocDefault));
    pin_bytes = bytes;
  }
};

namespace acu {
namespace {

// ---- host tables: the reference's FP32 expressions, one rounding each ----
#pragma clang fp contract(off)
float ref_pi() { return 4.0f * atanf(1.0f); }

// tv_hpc.F90:53-56
void make_filter(int N, float* h) {
  const float pi = ref_pi();
  for (int jx = 1; jx <= 2 * N + 1; ++jx) {
    const int k = jx - 1 - N;
    const int q = 4 * k * k - 1;
    h[jx - 1] = (-(2.0f * (float)N)) / ((pi * pi) * (float)q);
  }
}

// tv_hpc.F90:58-74
void make_cmask(int N, int T, int wpu, float* c) {
  const float pi = ref_pi();
  const float rad_n = (float)(N - 1) / 2.0f - (float)wpu;
  const float half = ((float)N - 1.0f) / 2.0f;
  for (int jy = 1; jy <= N; ++jy)  // each pixel on its own: rows in memory order
    for (int jx = 1; jx <= N; ++jx) {
      const float dx = ((float)jx - half) - 1.0f;
      const float dy = ((float)jy - half) - 1.0f;
      const float rr = sqrtf(dx * dx + dy * dy);
      const float rtmp = 1.0f - (2.0f * (rr - rad_n)) / (float)wpu;
      float v;
      if (rtmp <= 0.0f) v = 0.0f;
      else if (rtmp >= 1.0f) v = pi / (float)T;
      else {
        const float e1 = expf(1.0f - 1.0f / (1.0f - rtmp));
        v = (pi * expf(1.0f - 1.0f / (1.0f - e1))) / (float)T;
      }
      c[(size_t)(jy - 1) * N + (jx - 1)] = v;
    }
}

// ctheta of backprojection (tv_hpc.F90:113) or projection (:142), sin/cos once per angle
void make_angles(int T, const float* th, bool bp, float2* sc) {
  const float pi = ref_pi();
  for (int j = 0; j < T; ++j) {
    const float t = th[j];
    const float ct = bp ? (pi * (90.0f - t)) / 180.0f : (pi * (t - 90.0f)) / 180.0f;
    sc[j] = make_float2(sinf(ct), cosf(ct));
  }
}
#pragma clang fp contract(on)

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Carves typed arrays out of one device allocation (two passes: measure, then place).
struct Carver {
  char* base = nullptr;
  size_t off = 0;
  template <class T> T* take(size_t n) {
    off = align_up(off, 256);
    T* p = reinterpret_cast<T*>(base + off);
    off += n * sizeof(T);
    return p;
  }
};

template <class T>
struct DevArray {
  T* d = nullptr;
  int n = 0;
  // The copy goes on the context's stream and is waited for (h may be a
  // temporary): a copy on the legacy null stream fails while another
  // thread's stream captures a graph.
  void upload(const std::vector<T>& h, hipStream_t st) {
    if (d) { hipFree(d); d = nullptr; }
    n = (int)h.size();
    if (n == 0) return;
    ACU_HIP_CHECK(hipMalloc(&d, sizeof(T) * h.size()));
    ACU_HIP_CHECK(hipMemcpyAsync(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, st));
    ACU_HIP_CHECK(hipStreamSynchronize(st));
  }
  void release() { if (d) hipFree(d); d = nullptr; n = 0; }
};

struct Slice {
  int N = 0, T = 0, wpu = 0;
  float mid = 0.f;
  std::vector<float> theta;
  bool theta_set = false;
  int ntx = 0, nty = 0, nsub = 0;  // backprojection tiles along x, y; 16-row sub-tiles along y
  std::vector<uint8_t> tact_h;     // host copy of the sub-tile activity (tact)
  BPShares bp_sh{};                // XCD shares of this slice's own backprojection launch
  // operator arena
  float *h = nullptr, *cmask = nullptr;
  float2 *scbp = nullptr, *scfp = nullptr;
  float* rowterm = nullptr;  // backprojection row terms cos(c_t) * (y - mid), T*N + pad
  uint8_t* tact = nullptr;
  float *f = nullptr, *sinoCS = nullptr, *fbp = nullptr, *ruk = nullptr, *rp = nullptr;
  float4 *cells = nullptr, *cellsT = nullptr;
  float* fpart = nullptr;  // per-chunk partial ray sums of the forward projector
  int fp_clen = FP_CHUNK_DEFAULT;
  // forward-projection variant: 0 = global gathers of bilinear cells
  // (fp.hip), 1 / 2 / 4 = LDS boxes with that many angles per lane (fpl.hip)
  int fpv = 2;
  bool fpl_cells = true;  // the LDS projector may need the global cells of this slice
  uint8_t* pcls[3] = {nullptr, nullptr, nullptr};  // pitch classes of the angle blocks, al = 1, 2, 4
  // TV arena
  float *uk = nullptr, *ukp = nullptr, *dkx = nullptr, *dky = nullptr, *bkx = nullptr,
        *bky = nullptr, *ifkt = nullptr, *iukt = nullptr, *r = nullptr, *p = nullptr,
        *irpt = nullptr, *Ap = nullptr, *fk = nullptr;
  float* ftv = nullptr;  // iradon(f) with tvopt's backprojection positions, when they differ from FBP's
  double *ifkt64 = nullptr, *iukt64 = nullptr, *ruk64 = nullptr;
  double *scal = nullptr, *partials = nullptr;
  float* delta = nullptr;
  int delta_cap = 0;
  bool fbp_valid = false, tv_valid = false;
  // differentiation arena (allocated on first use): image, its gradient,
  // radon(x), shifted residual back-shifted, shifts, shift gradient,
  // data / TV partials and the three loss values
  float *gx_x = nullptr, *gx_g = nullptr, *gx_y = nullptr, *gx_rs = nullptr, *gx_sh = nullptr,
        *gx_gs = nullptr, *gx_hs = nullptr;
  double *gx_pd = nullptr, *gx_pt = nullptr, *gx_loss = nullptr;
  // global block ranges in the batched launches
  float* fftG = nullptr;   // FFT ramp: spectrum of the wrapped kernel / L (fft_len(N))
  float2* fftW = nullptr;  // FFT ramp: twiddles (fft_len(N) / 2)
  int ff_begin = 0, ff_n = 0;
  float ff_edge = 0.0f;    // FFT ramp: the tap at distance N-1 (h(2N-1), 0-based)
  int rf_begin = 0, rf_n = 0, bp_begin = 0, bp_n = 0, cl_begin = 0, cl_n = 0,
      fp_begin = 0, fp_n = 0, fs_begin = 0, fs_n = 0, fl_begin = 0, fl_n = 0, ew_begin = 0, ew_n = 0, sw_begin = 0, sw_n = 0;
};

// Ramp-filter task arrays of one operand pair: the Toeplitz GEMM and the FFT form.
struct RFSet {
  DevArray<RFTask> rf;
  DevArray<FFTTask> ff;
  void release() { rf.release(); ff.release(); }
};

// Task arrays of one radon operand pair, one array per variant group; a
// slice appears in the cells array only when its variant reads the cells.
struct FPSet {
  DevArray<CellTask> cl;
  DevArray<FPLTask> fl[4];  // one per LDS variant group (FL_GROUPS)
  DevArray<FPTask> fp;
  void release() {
    cl.release();
    for (auto& a : fl) a.release();
    fp.release();
  }
};

// task group of an LDS projector variant: al = 1, 2, 4, FPV_CELL -> 0..3
inline int fl_group(int al) { return al == FPV_CELL ? 3 : al >= 4 ? 2 : al == 2 ? 1 : 0; }
constexpr int FL_GROUPS = 4;
static_assert(sizeof(FPSet::fl) / sizeof(FPSet::fl[0]) == FL_GROUPS, "one task array per variant group");
inline bool fpv_valid(int v) { return v == 0 || v == 1 || v == 2 || v == 4 || v == FPV_CELL; }

}  // namespace
}  // namespace acu

struct acu_batch {
  acu_ctx* ctx = nullptr;
  std::vector<acu::Slice> s;
  void* ops_mem = nullptr;
  void* tv_mem = nullptr;
  int tv_delta_cap = 0;
  int ff_total = 0, ff_maxL = 0;  // FFT ramp blocks and the largest transform
  int rf_total = 0, bp_total = 0, cl_total = 0, fp_total = 0, fs_total = 0, ew_total = 0,
      sw_total = 0;
  int fl_total[acu::FL_GROUPS] = {0, 0, 0, 0};  // LDS projector blocks per variant group (lds1, lds2, lds4, cell16)
  // dispatch orders of the LDS projector (fpl_dispatch_order) per variant
  // group: of the launch over all slices, and of each slice's own launch at
  // [fl_begin, fl_begin + fl_n); empty when ACUMEN_FPL_ORDER=rb
  acu::DevArray<int> fl_ord_all[acu::FL_GROUPS], fl_ord_one[acu::FL_GROUPS];
  bool fpl_order = true;
  int ramp_variant = acu::RAMP_FP64;  // arithmetic of the ramp filter
  int bp_ty = acu::BP_TY_DEFAULT;     // backprojection tile height (16, 32 or 64)
  acu::BPShares bp_sh_all{};          // XCD shares of the backprojection over all slices
  bool bp_refpos = true;              // FBP backprojection positions in the reference's arithmetic
  bool tv_bp_refpos = true;           // the same inside tvopt's iteration
  // task sets for the fixed operand combinations of FBP / TV
  acu::RFSet rf_f, rf_fk, rf_ruk, rf_rp;  // ramp filter of f, fk, ruk, rp
  acu::DevArray<acu::BPTask> bp_fbp, bp_ftv, bp_ifkt, bp_iukt, bp_irpt;
  acu::FPSet fp_uk, fp_p;  // radon(uk) -> ruk, radon(p) -> rp
  acu::DevArray<acu::TVTask> tv;
  float tv_mu = 0.f, tv_lambda = 0.f;
  // single-slot task buffer for calls on caller-provided device pointers
  void* adhoc = nullptr;
  // differentiation: arena and per-call task sets
  void* grad_mem = nullptr;
  acu::FPSet fp_g;
  acu::DevArray<acu::ResTask> res_g;
  acu::DevArray<acu::AdjTask> adj_g;
  acu::DevArray<acu::TVGTask> tvg_g;
  acu::DevArray<acu::LossTask> loss_g;
  // operator schedule of an outer iteration (ACU_TV_SCHED_*) and, for the
  // linear schedule, the exact refresh period in outer iterations (0: none)
  int tv_sched = ACU_TV_SCHED_LINEAR;
  int tv_refresh = 0;
  // last tvopt's outer-iteration count (the length of its delta log)
  int last_nouter = -1;
  // graphs of one outer iteration: [0] the regular body, [1] the linear
  // schedule's refresh body
  hipGraphExec_t graph = nullptr, graph_refresh = nullptr;
  int graph_ninner = -1, graph_ncg = -1, graph_sched = -1;

  void drop_graphs() {
    if (graph) (void)hipGraphExecDestroy(graph);
    if (graph_refresh) (void)hipGraphExecDestroy(graph_refresh);
    graph = graph_refresh = nullptr;
  }

  ~acu_batch() {
    drop_graphs();
    rf_f.release(); rf_fk.release(); rf_ruk.release(); rf_rp.release();
    bp_fbp.release(); bp_ftv.release(); bp_ifkt.release(); bp_iukt.release(); bp_irpt.release();
    fp_uk.release(); fp_p.release(); tv.release();
    for (auto& o : fl_ord_all) o.release();
    for (auto& o : fl_ord_one) o.release();
    fp_g.release(); res_g.release(); adj_g.release();
    tvg_g.release(); loss_g.release();
    if (grad_mem) hipFree(grad_mem);
    if (adhoc) hipFree(adhoc);
    if (tv_mem) hipFree(tv_mem);
    if (ops_mem) hipFree(ops_mem);
  }
};

namespace acu {
namespace {

enum Kind { K_RF = 0, K_BP = 1, K_CELLS = 2, K_FP = 3, K_ADJ = 4, K_PREP = 5, K_INGEST = 6 };

template <class F>
void timed(acu_ctx* c, int kind, F&& launch) {
  if (!c->timing) { launch(); return; }
  hipEvent_t a = c->ev(), b = c->ev();
  ACU_HIP_CHECK(hipEventRecord(a, c->stream));
  launch();
  ACU_HIP_CHECK(hipEventRecord(b, c->stream));
  c->pending.push_back({kind, a, b});
}

void check_launch() { ACU_HIP_CHECK(hipGetLastError()); }

// Operator arena of every slice: tables, f, filtered sinogram, FBP image, cells.
void layout_ops(acu_batch* b, Carver& cv) {
  for (auto& s : b->s) {
    const size_t NN = (size_t)s.N * s.N, NT = (size_t)s.N * s.T, P = (size_t)(s.N + 1);
    s.h = cv.take<float>(2 * s.N + 1);
    s.cmask = cv.take<float>(NN);
    s.scbp = cv.take<float2>(s.T);
    s.scfp = cv.take<float2>(s.T);
    s.rowterm = cv.take<float>(NT + BP_ROWTERM_PAD);
    s.tact = cv.take<uint8_t>((size_t)s.ntx * s.nsub);
    s.f = cv.take<float>(NT);
    s.sinoCS = cv.take<float>(NT);
    s.fbp = cv.take<float>(NN);
    s.ruk = cv.take<float>(NT);
    s.rp = cv.take<float>(NT);
    s.cells = cv.take<float4>(P * P);
    s.cellsT = cv.take<float4>(P * P);
    s.fpart = cv.take<float>((size_t)fp_chunks(s.N, s.fp_clen) * NT);
    for (int g = 0; g < 3; ++g) s.pcls[g] = cv.take<uint8_t>((size_t)(s.T + (4 << g) - 1) / (4 << g));
    if (fft_supported(s.N)) {
      s.fftG = cv.take<float>((size_t)fft_len(s.N));
      s.fftW = cv.take<float2>((size_t)fft_len(s.N) / 2);
    }
  }
}

void layout_tv(acu_batch* b, Carver& cv, int delta_cap) {
  for (auto& s : b->s) {
    const size_t NN = (size_t)s.N * s.N, NT = (size_t)s.N * s.T;
    float** imgs[] = {&s.uk, &s.ukp, &s.dkx, &s.dky, &s.bkx, &s.bky,
                      &s.ifkt, &s.iukt, &s.r, &s.p, &s.irpt, &s.Ap};
    for (float** a : imgs) *a = cv.take<float>(NN);
    s.ftv = cv.take<float>(NN);
    s.fk = cv.take<float>(NT);
    // the linear schedule's carried terms, FP64
    s.ifkt64 = cv.take<double>(NN);
    s.iukt64 = cv.take<double>(NN);
    s.ruk64 = cv.take<double>(NT);
    s.scal = cv.take<double>(SCAL_STRIDE);
    s.partials = cv.take<double>((size_t)s.ew_n);
    s.delta = cv.take<float>((size_t)delta_cap);
    s.delta_cap = delta_cap;
  }
}

void upload_tables(acu_batch* b, int i) {
  auto& s = b->s[i];
  std::vector<float> h(2 * s.N + 1), cm((size_t)s.N * s.N);
  make_filter(s.N, h.data());
  make_cmask(s.N, s.T, s.wpu, cm.data());
  std::vector<uint8_t> act((size_t)s.ntx * s.nsub, 0);
  for (int y = 0; y < s.N; ++y)
    for (int x = 0; x < s.N; ++x)
      if (cm[(size_t)y * s.N + x] != 0.0f) act[(size_t)(y / BP_SUB) * s.ntx + x / BP_TX] = 1;
  hipStream_t st = b->ctx->stream;
  ACU_HIP_CHECK(hipMemcpyAsync(s.h, h.data(), h.size() * 4, hipMemcpyHostToDevice, st));
  s.ff_edge = h[2 * s.N - 1];
  if (s.fftG) {
    const int L = fft_len(s.N);
    std::vector<float> G(L);
    std::vector<float2> W(L / 2);
    fft_tables(s.N, h.data(), G.data(), W.data());
    ACU_HIP_CHECK(hipMemcpyAsync(s.fftG, G.data(), G.size() * 4, hipMemcpyHostToDevice, st));
    ACU_HIP_CHECK(hipMemcpyAsync(s.fftW, W.data(), W.size() * sizeof(float2), hipMemcpyHostToDevice, st));
    ACU_HIP_CHECK(hipStreamSynchronize(st));  // G and W end here
  }
  ACU_HIP_CHECK(hipMemcpyAsync(s.cmask, cm.data(), cm.size() * 4, hipMemcpyHostToDevice, st));
  ACU_HIP_CHECK(hipMemcpyAsync(s.tact, act.data(), act.size(), hipMemcpyHostToDevice, st));
  ACU_HIP_CHECK(hipStreamSynchronize(st));
  s.tact_h = std::move(act);
}

// Tables of slice i from slice j of the same (N, T, wpu), device to device:
// a batch builds its host tables once per distinct geometry.
void copy_tables(acu_batch* b, int j, int i) {
  const auto& a = b->s[j];
  auto& s = b->s[i];
  hipStream_t st = b->ctx->stream;  // ordered before every later use of the tables
  ACU_HIP_CHECK(hipMemcpyAsync(s.h, a.h, (size_t)(2 * s.N + 1) * 4, hipMemcpyDeviceToDevice, st));
  ACU_HIP_CHECK(hipMemcpyAsync(s.cmask, a.cmask, (size_t)s.N * s.N * 4, hipMemcpyDeviceToDevice, st));
  ACU_HIP_CHECK(hipMemcpyAsync(s.tact, a.tact, (size_t)s.ntx * s.nsub, hipMemcpyDeviceToDevice, st));
  s.tact_h = a.tact_h;
  s.ff_edge = a.ff_edge;
  if (s.fftG) {
    const size_t L = (size_t)fft_len(s.N);
    ACU_HIP_CHECK(hipMemcpyAsync(s.fftG, a.fftG, L * 4, hipMemcpyDeviceToDevice, st));
    ACU_HIP_CHECK(hipMemcpyAsync(s.fftW, a.fftW, L / 2 * sizeof(float2), hipMemcpyDeviceToDevice, st));
  }
}

uint64_t theta_hash(const std::vector<float>& th);

// Pitch classes of a slice geometry (host model of the LDS banks, a few tens
// of ms per geometry): computed once per (N, T, angles, angles per lane) in
// the process and shared by every slice and batch of that geometry.
const std::vector<uint8_t>& pitch_classes_cached(int N, int T, const std::vector<float>& theta, int al) {
  static std::mutex mu;
  static std::map<std::tuple<int, int, uint64_t, int>, std::vector<uint8_t>> cache;
  const auto key = std::make_tuple(N, T, theta_hash(theta), al);
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find(key);
  if (it == cache.end()) {
    std::vector<uint8_t> pc((size_t)(T + 4 * al - 1) / (4 * al));
    fpl_pitch_classes(N, T, theta.data(), al, pc.data());
    it = cache.emplace(key, std::move(pc)).first;
  }
  return it->second;
}

void upload_theta(acu_batch* b, int i) {
  auto& s = b->s[i];
  std::vector<float2> bp(s.T), fp(s.T);
  make_angles(s.T, s.theta.data(), true, bp.data());
  make_angles(s.T, s.theta.data(), false, fp.data());
  hipStream_t st = b->ctx->stream;
  ACU_HIP_CHECK(hipMemcpyAsync(s.scbp, bp.data(), bp.size() * sizeof(float2), hipMemcpyHostToDevice, st));
  ACU_HIP_CHECK(hipMemcpyAsync(s.scfp, fp.data(), fp.size() * sizeof(float2), hipMemcpyHostToDevice, st));
  for (int g = 0; g < 3; ++g) {
    const std::vector<uint8_t>& pc = pitch_classes_cached(s.N, s.T, s.theta, 1 << g);
    ACU_HIP_CHECK(hipMemcpyAsync(s.pcls[g], pc.data(), pc.size(), hipMemcpyHostToDevice, st));
  }
  launch_bp_rowterm(s.scbp, s.N, s.T, s.mid, s.rowterm, st);
  ACU_HIP_CHECK(hipGetLastError());
  ACU_HIP_CHECK(hipStreamSynchronize(st));  // bp and fp end here
  s.theta_set = true;
}

RFTask rf_task(const Slice& s, const float* in, float* out) {
  return RFTask{s.N, s.T, s.rf_begin, rf_tiles_m(s.N), rf_tiles_n(s.T), s.h, in, out};
}
FFTTask ff_task(const Slice& s, const float* in, float* out) {
  const int L = fft_len(s.N);
  int lg = 0;
  while ((1 << lg) < L) ++lg;
  return FFTTask{s.N, s.T, L, lg, s.ff_begin, L == 2 * s.N - 2 ? 1 : 0, s.ff_edge, s.fftG, s.fftW, in, out};
}
BPTask bp_task(const Slice& s, const float* in, float* out) {
  return BPTask{s.N, s.T, s.mid, s.bp_begin, s.ntx, s.nty, s.nsub, in, s.cmask, s.scbp, s.tact, s.rowterm, out};
}
FPLTask fpl_task(const Slice& s, const float* img, float* out) {
  static const bool fixed_pitch = [] {
    const char* e = std::getenv("ACUMEN_FPL_PITCH");
    return e && std::strcmp(e, "fixed") == 0;
  }();
  const int blk = fpl_angle_block(s.fpv);
  const uint8_t* pc = (fixed_pitch || s.fpv == FPV_CELL) ? nullptr : s.pcls[fl_group(s.fpv)];
  return FPLTask{s.N, s.T, s.mid, s.fl_begin, s.fpv, (s.T + blk - 1) / blk, img, s.cells, s.scfp, out, pc};
}
CellTask cell_task(const Slice& s, const float* img) {
  return CellTask{s.N, s.cl_begin, img, s.cells, s.cellsT};
}
FPTask fp_task(const Slice& s, float* out) {
  return FPTask{s.N, s.T, s.mid, s.fp_begin, fp_ray_blocks(s.N), fp_chunks(s.N, s.fp_clen), s.fp_clen,
                s.fs_begin, s.fpart, s.cells, s.cellsT, s.scfp, out};
}

template <class T, class F>
void build(acu_batch* b, DevArray<T>& arr, F&& make) {
  std::vector<T> v;
  for (auto& s : b->s) v.push_back(make(s));
  arr.upload(v, b->ctx->stream);
}

// Block ranges of the forward-projection groups from each slice's variant:
// cells only for slices whose variant reads them, LDS-projector blocks
// within the slice's angles-per-lane group, gather blocks within the
// gather group.
void assign_fp(acu_batch* b) {
  int cl = 0, fp = 0, fs = 0, fl[FL_GROUPS] = {0, 0, 0, 0};
  for (auto& s : b->s) {
    s.fpl_cells = s.fpv == 0 || !s.theta_set || fpl_may_overflow(s.N, s.T, s.theta.data(), s.fpv);
    s.cl_begin = cl; s.cl_n = s.fpl_cells ? cell_blocks(s.N) : 0; cl += s.cl_n;
    s.fl_begin = 0; s.fl_n = 0; s.fp_begin = fp; s.fp_n = 0; s.fs_begin = fs; s.fs_n = 0;
    if (s.fpv > 0) {
      const int g = fl_group(s.fpv);
      s.fl_begin = fl[g]; s.fl_n = fpl_blocks(s.N, s.T, s.fpv); fl[g] += s.fl_n;
    } else {
      s.fp_n = fp_ray_blocks(s.N) * s.T * fp_chunks(s.N, s.fp_clen); fp += s.fp_n;
      s.fs_n = fp_sum_blocks(s.N, s.T); fs += s.fs_n;
    }
  }
  b->cl_total = cl; b->fp_total = fp; b->fs_total = fs;
  for (int g = 0; g < FL_GROUPS; ++g) b->fl_total[g] = fl[g];
  // dispatch orders (the cell variant keeps its own order)
  for (int g = 0; g < FL_GROUPS; ++g) {
    b->fl_ord_all[g].release();
    b->fl_ord_one[g].release();
    if (!b->fpl_order || g == fl_group(FPV_CELL) || fl[g] == 0) continue;
    std::vector<FplGeom> geo;
    std::vector<int> one((size_t)fl[g]);
    for (const auto& s : b->s) {
      if (s.fpv <= 0 || fl_group(s.fpv) != g) continue;
      const FplGeom t{s.N, (s.T + fpl_angle_block(s.fpv) - 1) / fpl_angle_block(s.fpv), s.fl_begin};
      geo.push_back(t);
      const std::vector<int> o = fpl_dispatch_order({t}, s.fl_begin, s.fl_n);
      if (o.empty())  // the natural order
        for (int i = 0; i < s.fl_n; ++i) one[(size_t)s.fl_begin + i] = i;
      else
        std::copy(o.begin(), o.end(), one.begin() + s.fl_begin);
    }
    b->fl_ord_all[g].upload(fpl_dispatch_order(geo, 0, fl[g]), b->ctx->stream);
    b->fl_ord_one[g].upload(one, b->ctx->stream);
  }
}

// dispatch order of an LDS projector launch (null: ray blocks in order)
const int* fl_order_all(const acu_batch* b, int g) { return b->fl_ord_all[g].d; }
const int* fl_order_one(const acu_batch* b, const Slice& s) {
  const auto& o = b->fl_ord_one[fl_group(s.fpv)];
  return o.d ? o.d + s.fl_begin : nullptr;
}

template <class ImgF, class OutF>
void build_fpset(acu_batch* b, FPSet& set, ImgF img, OutF out) {
  std::vector<CellTask> cl;
  std::vector<FPLTask> fl[FL_GROUPS];
  std::vector<FPTask> fp;
  for (auto& s : b->s) {
    if (s.cl_n > 0) cl.push_back(cell_task(s, img(s)));
    if (s.fpv > 0) fl[fl_group(s.fpv)].push_back(fpl_task(s, img(s), out(s)));
    else fp.push_back(fp_task(s, out(s)));
  }
  set.cl.upload(cl, b->ctx->stream);
  for (int g = 0; g < FL_GROUPS; ++g) set.fl[g].upload(fl[g], b->ctx->stream);
  set.fp.upload(fp, b->ctx->stream);
}

void build_tv_tasks(acu_batch* b);

template <class InF, class OutF>
void build_rfset(acu_batch* b, RFSet& set, InF in, OutF out) {
  std::vector<RFTask> rf;
  std::vector<FFTTask> ff;
  for (auto& s : b->s) {
    rf.push_back(rf_task(s, in(s), out(s)));
    if (s.fftG) ff.push_back(ff_task(s, in(s), out(s)));
  }
  set.rf.upload(rf, b->ctx->stream);
  set.ff.upload(ff, b->ctx->stream);
}

// After a change of variants or angles: new ranges, new task arrays, and a
// captured outer iteration is rebuilt on the next tvopt.
void refresh_fp(acu_batch* b) {
  assign_fp(b);
  if (b->tv_mem) build_tv_tasks(b);
  b->drop_graphs();
}

// Estimated work of slice s's backprojection tiles at height ty, in launch
// order: T for a tile with a mask pixel, 2% of that for one that only stores
// zeros (every slice's tables must be built).
void bp_tile_work(const Slice& s, int ty, std::vector<float>& w) {
  const int nty = bp_tiles_y(s.N, ty), subs = ty / BP_SUB;
  for (int y = 0; y < nty; ++y)
    for (int x = 0; x < s.ntx; ++x) {
      bool act = false;
      for (int j = 0; j < subs; ++j) {
        const int r = y * subs + j;
        if (r < s.nsub) act = act || s.tact_h[(size_t)r * s.ntx + x] != 0;
      }
      w.push_back(act ? (float)s.T : 0.02f * (float)s.T);
    }
}

// XCD shares of the all-slice and per-slice backprojection launches.
void assign_bp_shares(acu_batch* b) {
  std::vector<float> all, one;
  for (auto& s : b->s) {
    one.clear();
    bp_tile_work(s, b->bp_ty, one);
    s.bp_sh = bp_xcd_shares(one);
    all.insert(all.end(), one.begin(), one.end());
  }
  b->bp_sh_all = bp_xcd_shares(all);
}

// Backprojection tile ranges at the batch's tile height (and their XCD
// shares once the slices' tables exist).
void assign_bp(acu_batch* b) {
  int bp = 0;
  bool tables = true;
  for (auto& s : b->s) {
    s.nty = bp_tiles_y(s.N, b->bp_ty);
    s.bp_begin = bp; s.bp_n = s.ntx * s.nty; bp += s.bp_n;
    tables = tables && !s.tact_h.empty();
  }
  b->bp_total = bp;
  if (tables) assign_bp_shares(b);
}

void build_op_tasks(acu_batch* b);

// A new backprojection tile height: new ranges and task arrays, and a
// captured outer iteration is rebuilt on the next tvopt.
void set_bp_tile(acu_batch* b, int ty) {
  b->bp_ty = ty;
  assign_bp(b);
  build_op_tasks(b);
  if (b->tv_mem) build_tv_tasks(b);
  b->drop_graphs();
}

void build_op_tasks(acu_batch* b) {
  build_rfset(b, b->rf_f, [](Slice& s) { return (const float*)s.f; }, [](Slice& s) { return s.sinoCS; });
  build(b, b->bp_fbp, [](Slice& s) { return bp_task(s, s.sinoCS, s.fbp); });
}

void build_tv_tasks(acu_batch* b) {
  build_rfset(b, b->rf_fk, [](Slice& s) { return (const float*)s.fk; }, [](Slice& s) { return s.sinoCS; });
  build_rfset(b, b->rf_ruk, [](Slice& s) { return (const float*)s.ruk; }, [](Slice& s) { return s.sinoCS; });
  build_rfset(b, b->rf_rp, [](Slice& s) { return (const float*)s.rp; }, [](Slice& s) { return s.sinoCS; });
  build(b, b->bp_ftv, [](Slice& s) { return bp_task(s, s.sinoCS, s.ftv); });
  build(b, b->bp_ifkt, [](Slice& s) { return bp_task(s, s.sinoCS, s.ifkt); });
  build(b, b->bp_iukt, [](Slice& s) { return bp_task(s, s.sinoCS, s.iukt); });
  build(b, b->bp_irpt, [](Slice& s) { return bp_task(s, s.sinoCS, s.irpt); });
  build_fpset(b, b->fp_uk, [](Slice& s) { return (const float*)s.uk; }, [](Slice& s) { return s.ruk; });
  build_fpset(b, b->fp_p, [](Slice& s) { return (const float*)s.p; }, [](Slice& s) { return s.rp; });
}

// tvopt's backprojection positions differ from the FBP's: its iteration
// carries iradon(f) with its own positions (s.ftv) instead of the FBP image
bool tv_own_fbp(const acu_batch* b) { return b->tv_bp_refpos != b->bp_refpos; }

void build_tv_task_array(acu_batch* b) {
  build(b, b->tv, [b](Slice& s) {
    TVTask t{};
    t.N = s.N; t.T = s.T;
    t.eb_begin = s.ew_begin; t.nb = s.ew_n;
    t.sb_begin = s.sw_begin; t.nsb = s.sw_n;
    t.mu = b->tv_mu; t.lambda = b->tv_lambda;
    t.uk = s.uk; t.ukp = s.ukp; t.dkx = s.dkx; t.dky = s.dky; t.bkx = s.bkx; t.bky = s.bky;
    t.ifkt = s.ifkt; t.iukt = s.iukt; t.r = s.r; t.p = s.p; t.irpt = s.irpt; t.Ap = s.Ap;
    t.f = s.f; t.fk = s.fk; t.ruk = s.ruk; t.rp = s.rp; t.fbp = tv_own_fbp(b) ? s.ftv : s.fbp;
    t.ifkt64 = s.ifkt64; t.iukt64 = s.iukt64; t.ruk64 = s.ruk64;
    t.scal = s.scal; t.partials = s.partials;
    t.delta_log = s.delta; t.delta_cap = s.delta_cap;
    return t;
  });
}

void ensure_tv(acu_batch* b, int delta_cap) {
  if (b->tv_mem && delta_cap <= b->tv_delta_cap) return;
  if (b->tv_mem) { hipFree(b->tv_mem); b->tv_mem = nullptr; }
  b->drop_graphs();
  Carver dry;
  layout_tv(b, dry, delta_cap);
  ACU_HIP_CHECK(hipMalloc(&b->tv_mem, dry.off + 256));
  Carver cv{static_cast<char*>(b->tv_mem), 0};
  layout_tv(b, cv, delta_cap);
  for (auto& s : b->s)
    ACU_HIP_CHECK(hipMemsetAsync(s.scal, 0, SCAL_STRIDE * sizeof(double), b->ctx->stream));
  b->tv_delta_cap = delta_cap;
  build_tv_tasks(b);
  // the fused TV tasks point into the old layout: tvopt rebuilds them
  b->tv.release();
}

void require_theta(acu_batch* b) {
  for (size_t i = 0; i < b->s.size(); ++i)
    if (!b->s[i].theta_set) throw Error("theta of slice " + std::to_string(i) + " was never set");
}

// ---- operator sequences (all slices at once, or one slice: sl >= 0) ----
struct Range { int base, n; };
Range rf_range(acu_batch* b, int sl) { return sl < 0 ? Range{0, b->rf_total} : Range{b->s[sl].rf_begin, b->s[sl].rf_n}; }
Range bp_range(acu_batch* b, int sl) { return sl < 0 ? Range{0, b->bp_total} : Range{b->s[sl].bp_begin, b->s[sl].bp_n}; }
const BPShares& bp_shares(const acu_batch* b, int sl) { return sl < 0 ? b->bp_sh_all : b->s[sl].bp_sh; }
Range ew_range(acu_batch* b, int sl) { return sl < 0 ? Range{0, b->ew_total} : Range{b->s[sl].ew_begin, b->s[sl].ew_n}; }
Range sw_range(acu_batch* b, int sl) { return sl < 0 ? Range{0, b->sw_total} : Range{b->s[sl].sw_begin, b->s[sl].sw_n}; }

void ramp_seq(acu_batch* b, const RFSet& rf, int sl) {
  acu_ctx* c = b->ctx;
  if (b->ramp_variant == RAMP_FFT) {
    const Range r = sl < 0 ? Range{0, b->ff_total} : Range{b->s[sl].ff_begin, b->s[sl].ff_n};
    timed(c, K_RF, [&] { launch_ramp_fft(rf.ff.d, rf.ff.n, r.base, r.n, b->ff_maxL, c->stream); });
  } else {
    const Range r = rf_range(b, sl);
    timed(c, K_RF, [&] { launch_ramp(rf.rf.d, rf.rf.n, r.base, r.n, b->ramp_variant, c->stream); });
  }
}

// refpos: the backprojection positions (the batch's FBP positions, or
// tvopt's inside its iteration)
void iradon_seq(acu_batch* b, const RFSet& rf, const DevArray<BPTask>& bp, int sl, bool refpos) {
  acu_ctx* c = b->ctx;
  const Range r2 = bp_range(b, sl);
  ramp_seq(b, rf, sl);
  timed(c, K_BP, [&] { launch_backproject(bp.d, bp.n, r2.base, r2.n, b->bp_ty, refpos, bp_shares(b, sl), c->stream); });
  check_launch();
}

// radon of all slices (sl < 0) or of one slice with the task arrays of one
// operand pair: cells where a slice's variant reads them, then one launch
// per non-empty variant group.
void radon_seq(acu_batch* b, const FPSet& set, int sl) {
  acu_ctx* c = b->ctx;
  static const int al_of[FL_GROUPS] = {1, 2, 4, FPV_CELL};
  if (sl < 0) {
    if (b->cl_total > 0)
      timed(c, K_CELLS, [&] { launch_cells(set.cl.d, set.cl.n, 0, b->cl_total, c->stream); });
    for (int g = 0; g < FL_GROUPS; ++g)
      if (b->fl_total[g] > 0)
        timed(c, K_FP, [&] {
          launch_forward_lds(set.fl[g].d, set.fl[g].n, 0, b->fl_total[g], al_of[g], fl_order_all(b, g),
                             c->stream);
        });
    if (b->fp_total > 0)
      timed(c, K_FP, [&] { launch_forward(set.fp.d, set.fp.n, 0, b->fp_total, 0, b->fs_total, c->stream); });
  } else {
    const Slice& s = b->s[sl];
    if (s.cl_n > 0)
      timed(c, K_CELLS, [&] { launch_cells(set.cl.d, set.cl.n, s.cl_begin, s.cl_n, c->stream); });
    if (s.fpv > 0) {
      const int g = fl_group(s.fpv);
      timed(c, K_FP, [&] {
        launch_forward_lds(set.fl[g].d, set.fl[g].n, s.fl_begin, s.fl_n, s.fpv, fl_order_one(b, s), c->stream);
      });
    } else {
      timed(c, K_FP, [&] {
        launch_forward(set.fp.d, set.fp.n, s.fp_begin, s.fp_n, s.fs_begin, s.fs_n, c->stream);
      });
    }
  }
  check_launch();
}

double read_scalar(acu_batch* b, int sl, int idx) {
  double v = 0.0;
  ACU_HIP_CHECK(hipMemcpyAsync(&v, b->s[sl].scal + idx, sizeof(double), hipMemcpyDeviceToHost,
                               b->ctx->stream));
  ACU_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
  return v;
}

// One outer iteration (tv_hpc.F90:205-353) for all slices or one slice.
// Operator schedule (b->tv_sched), per outer iteration with ninner = 1:
//  * REFERENCE: the reference's 7 radon + 7 iradon (:215, :228, :229, five
//    CG pairs, :339).
//  * REUSE: 6 + 7.  ruk_fresh says ruk already holds radon(uk) for the
//    current uk (after the prologue and after every add-back, :339), so the
//    radon of :228 in the first inner iteration is that same result, bit
//    for bit.
//  * LINEAR: 5 + 5.  R and F are linear, and the CG already forms
//    rp_j = R p_j and irpt_j = F R p_j, so with uk += sum_j alpha_j p_j:
//      ruk  = R uk   is carried as ruk  += alpha_j rp_j    (replaces :339, :228),
//      iukt = F R uk is carried as iukt += alpha_j irpt_j  (replaces :229),
//      ifkt = F fk   is carried as ifkt  = (ifkt + F f) - iukt after the
//                    add-back fk = (fk + f) - ruk          (replaces :215),
//    from exact values formed in the prologue.  The three are carried in
//    FP64 (ruk64, iukt64, ifkt64): carried in FP32, their rounding drifted
//    to 2e-4 from the FP64 result after 50 outer iterations at 513^2 x 360
//    (5.7e-5 in FP64, as the reference's own FP32).  refresh: recompute the
//    three exactly at the start of this outer iteration.
// host_sync: return every reduction to the host before the next kernel, as
// the reference does (the plain per-slice path).
void outer_body(acu_batch* b, int sl, int ninner, int ncg, bool ruk_fresh, bool host_sync, bool refresh) {
  acu_ctx* c = b->ctx;
  const Range ew = ew_range(b, sl), sw = sw_range(b, sl);
  const TVTask* tv = b->tv.d;
  const int nt = b->tv.n;
  const int first = sl < 0 ? 0 : sl, count = sl < 0 ? nt : 1;  // finalize: one block per slice
  const bool lin = b->tv_sched == ACU_TV_SCHED_LINEAR;
  for (int ji = 0; ji < ninner; ++ji) {
    const bool exact = !lin || (refresh && ji == 0);
    const bool tp = b->tv_bp_refpos;
    if (exact) iradon_seq(b, b->rf_fk, b->bp_ifkt, sl, tp);        // :215
    const bool reuse = b->tv_sched == ACU_TV_SCHED_REUSE && ruk_fresh && ji == 0;
    if (exact && !reuse) radon_seq(b, b->fp_uk, sl);               // :228
    if (exact) iradon_seq(b, b->rf_ruk, b->bp_iukt, sl, tp);       // :229
    if (lin && exact) launch_tv_linload(tv, nt, ew.base, ew.n, sw.base, sw.n, c->stream);
    launch_tv_residual(tv, nt, ew.base, ew.n, lin, c->stream);     // :207-249
    launch_tv_finalize(tv, first, count, 0, c->stream);
    check_launch();
    if (host_sync) (void)read_scalar(b, sl, 0);
    for (int j = 0; j < ncg; ++j) {                                // :252-303
      radon_seq(b, b->fp_p, sl);
      iradon_seq(b, b->rf_rp, b->bp_irpt, sl, tp);
      launch_tv_apdot(tv, nt, ew.base, ew.n, j, c->stream);
      launch_tv_finalize(tv, first, count, 64 + j, c->stream);
      check_launch();
      if (host_sync) (void)read_scalar(b, sl, 64 + j);
      launch_tv_axpy(tv, nt, ew.base, ew.n, j, lin, c->stream);
      if (lin) launch_tv_sino_axpy(tv, nt, sw.base, sw.n, j, c->stream);
      launch_tv_finalize(tv, first, count, j + 1, c->stream);
      check_launch();
      if (host_sync) (void)read_scalar(b, sl, j + 1);
      launch_tv_pupdate(tv, nt, ew.base, ew.n, j, c->stream);
      check_launch();
    }
    launch_tv_shrink(tv, nt, ew.base, ew.n, lin && ji == ninner - 1, c->stream);  // :305-333
    launch_tv_finalize(tv, first, count, 127, c->stream);
    check_launch();
    if (host_sync) (void)read_scalar(b, sl, 127);
  }
  if (!lin) radon_seq(b, b->fp_uk, sl);                            // :339
  launch_tv_addback(tv, nt, sw.base, sw.n, lin, c->stream);        // :341-346
  check_launch();
}

void fbp_all(acu_batch* b, int sl) {
  iradon_seq(b, b->rf_f, b->bp_fbp, sl, b->bp_refpos);
}

// the backprojection alone (its ramp-filtered input is already in place)
void bp_seq(acu_batch* b, const DevArray<BPTask>& bp, int sl, bool refpos) {
  acu_ctx* c = b->ctx;
  const Range r2 = bp_range(b, sl);
  timed(c, K_BP, [&] { launch_backproject(bp.d, bp.n, r2.base, r2.n, b->bp_ty, refpos, bp_shares(b, sl), c->stream); });
  check_launch();
}

// Prologue of tvopt (tv_hpc.F90:176-202) for all slices or one: FBP warm
// start, uk = fbp, d = b = 0, fk = f; then what the schedule carries:
// ruk = radon(uk) (REUSE, LINEAR), iukt = iradon(ruk) and ifkt = iradon(f)
// (LINEAR): the FBP itself, or, when tvopt's backprojection positions differ
// from the FBP's, the same ramp-filtered f backprojected with tvopt's
// positions into s.ftv.
void tv_prologue(acu_batch* b, int sl, int nouter) {
  acu_ctx* c = b->ctx;
  fbp_all(b, sl);
  const int i0 = sl < 0 ? 0 : sl, i1 = sl < 0 ? (int)b->s.size() : sl + 1;
  const bool lin = b->tv_sched == ACU_TV_SCHED_LINEAR;
  const bool own = lin && nouter > 0 && tv_own_fbp(b);
  if (own) bp_seq(b, b->bp_ftv, sl, b->tv_bp_refpos);  // sinoCS still holds the ramp of f
  for (int i = i0; i < i1; ++i) {
    const Slice& s = b->s[i];
    const size_t bytes = (size_t)s.N * s.N * 4;
    ACU_HIP_CHECK(hipMemcpyAsync(s.uk, s.fbp, bytes, hipMemcpyDeviceToDevice, c->stream));
    if (lin && nouter > 0)
      ACU_HIP_CHECK(hipMemcpyAsync(s.ifkt, own ? s.ftv : s.fbp, bytes, hipMemcpyDeviceToDevice, c->stream));
  }
  const Range ew = ew_range(b, sl), sw = sw_range(b, sl);
  launch_tv_init(b->tv.d, b->tv.n, ew.base, ew.n, sw.base, sw.n, c->stream);
  check_launch();
  if (nouter > 0 && b->tv_sched != ACU_TV_SCHED_REFERENCE) radon_seq(b, b->fp_uk, sl);
  if (nouter > 0 && lin) {
    iradon_seq(b, b->rf_ruk, b->bp_iukt, sl, b->tv_bp_refpos);
    launch_tv_linload(b->tv.d, b->tv.n, ew.base, ew.n, sw.base, sw.n, c->stream);
    check_launch();
  }
}

bool refresh_at(const acu_batch* b, int jo) {
  return b->tv_sched == ACU_TV_SCHED_LINEAR && b->tv_refresh > 0 && jo > 0 && jo % b->tv_refresh == 0;
}

hipGraphExec_t capture_outer(acu_batch* b, int ninner, int ncg, bool refresh) {
  acu_ctx* c = b->ctx;
  hipGraph_t g;
  ACU_HIP_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  try {
    outer_body(b, -1, ninner, ncg, true, false, refresh);
  } catch (...) {
    (void)hipStreamEndCapture(c->stream, &g);
    throw;
  }
  ACU_HIP_CHECK(hipStreamEndCapture(c->stream, &g));
  hipGraphExec_t e = nullptr;
  ACU_HIP_CHECK(hipGraphInstantiate(&e, g, nullptr, nullptr, 0));
  ACU_HIP_CHECK(hipGraphDestroy(g));
  return e;
}

void tvopt_run(acu_batch* b, float mu, float lambda, int nouter, int ninner, int ncg, int mode) {
  acu_ctx* c = b->ctx;
  require_theta(b);
  if (ncg < 0 || ncg > 60 || nouter < 0 || ninner < 0) throw Error("ncg must be in [0,60], nouter and ninner >= 0");
  ensure_tv(b, nouter > b->tv_delta_cap ? nouter : b->tv_delta_cap);
  if (mu != b->tv_mu || lambda != b->tv_lambda || b->tv.n == 0) {
    b->tv_mu = mu;
    b->tv_lambda = lambda;
    build_tv_task_array(b);
    b->drop_graphs();
  }
  for (auto& s : b->s) { s.fbp_valid = false; s.tv_valid = false; }
  b->last_nouter = -1;
  if (mode == ACU_TV_PLAIN) {
    // the reference's structure: one slice after another, every kernel of
    // the slice launched for it alone, every reduction back on the host
    for (int sl = 0; sl < (int)b->s.size(); ++sl) {
      tv_prologue(b, sl, nouter);
      for (int jo = 0; jo < nouter; ++jo) outer_body(b, sl, ninner, ncg, true, true, refresh_at(b, jo));
    }
  } else {
    tv_prologue(b, -1, nouter);
    const bool graph = (mode == ACU_TV_GRAPH) && !c->timing;
    if (!graph) {
      for (int jo = 0; jo < nouter; ++jo) outer_body(b, -1, ninner, ncg, true, false, refresh_at(b, jo));
    } else if (nouter > 0) {
      if (!b->graph || b->graph_ninner != ninner || b->graph_ncg != ncg || b->graph_sched != b->tv_sched) {
        b->drop_graphs();
        b->graph = capture_outer(b, ninner, ncg, false);
        b->graph_ninner = ninner;
        b->graph_ncg = ncg;
        b->graph_sched = b->tv_sched;
      }
      for (int jo = 0; jo < nouter; ++jo) {
        if (refresh_at(b, jo)) {
          if (!b->graph_refresh) b->graph_refresh = capture_outer(b, ninner, ncg, true);
          ACU_HIP_CHECK(hipGraphLaunch(b->graph_refresh, c->stream));
        } else {
          ACU_HIP_CHECK(hipGraphLaunch(b->graph, c->stream));
        }
      }
    }
  }
  for (auto& s : b->s) { s.fbp_valid = true; s.tv_valid = true; }
  b->last_nouter = nouter;
}

void adhoc_slot(acu_batch* b, size_t bytes) {
  if (!b->adhoc) ACU_HIP_CHECK(hipMalloc(&b->adhoc, 256));
  if (bytes > 256) throw Error("task too large");
  // the previous call on caller pointers may still read the slot
  ACU_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
}

template <class T>
T* put_adhoc(acu_batch* b, const T& t) {
  adhoc_slot(b, sizeof(T));
  ACU_HIP_CHECK(hipMemcpyAsync(b->adhoc, &t, sizeof(T), hipMemcpyHostToDevice, b->ctx->stream));
  ACU_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));  // t is the caller's temporary
  return static_cast<T*>(b->adhoc);
}

Slice& slice_of(acu_batch* b, int sl) {
  if (sl < 0 || sl >= (int)b->s.size()) throw Error("slice index out of range");
  if (!b->s[sl].theta_set) throw Error("theta of slice " + std::to_string(sl) + " was never set");
  return b->s[sl];
}

void radon_dev(acu_batch* b, int sl, float* sino, const float* img) {
  Slice& s = slice_of(b, sl);
  acu_ctx* c = b->ctx;
  if (s.cl_n > 0) {
    CellTask t = cell_task(s, img);
    t.block_begin = 0;
    CellTask* d = put_adhoc(b, t);
    timed(c, K_CELLS, [&] { launch_cells(d, 1, 0, s.cl_n, c->stream); });
  }
  if (s.fpv > 0) {
    FPLTask t = fpl_task(s, img, sino);
    t.block_begin = 0;
    FPLTask* d = put_adhoc(b, t);
    timed(c, K_FP, [&] { launch_forward_lds(d, 1, 0, s.fl_n, s.fpv, fl_order_one(b, s), c->stream); });
  } else {
    FPTask t = fp_task(s, sino);
    t.block_begin = 0;
    t.sum_begin = 0;
    FPTask* d = put_adhoc(b, t);
    timed(c, K_FP, [&] { launch_forward(d, 1, 0, s.fp_n, 0, s.fs_n, c->stream); });
  }
  check_launch();
}

void ramp_dev(acu_batch* b, int sl, float* out, const float* in) {
  Slice& s = slice_of(b, sl);
  acu_ctx* c = b->ctx;
  if (b->ramp_variant == RAMP_FFT) {
    FFTTask t = ff_task(s, in, out);
    t.block_begin = 0;
    FFTTask* d = put_adhoc(b, t);
    timed(c, K_RF, [&] { launch_ramp_fft(d, 1, 0, s.ff_n, fft_len(s.N), c->stream); });
    check_launch();
    return;
  }
  RFTask t = rf_task(s, in, out);
  t.tile_begin = 0;
  RFTask* d = put_adhoc(b, t);
  timed(c, K_RF, [&] { launch_ramp(d, 1, 0, s.rf_n, b->ramp_variant, c->stream); });
  check_launch();
}

void bp_dev(acu_batch* b, int sl, float* img, const float* filt) {
  Slice& s = slice_of(b, sl);
  acu_ctx* c = b->ctx;
  BPTask t = bp_task(s, filt, img);
  t.tile_begin = 0;
  BPTask* d = put_adhoc(b, t);
  timed(c, K_BP, [&] { launch_backproject(d, 1, 0, s.bp_n, b->bp_ty, b->bp_refpos, s.bp_sh, c->stream); });
  check_launch();
}

// ---- autotuner: the forward-projection variant of every slice, by measurement ----
uint64_t theta_hash(const std::vector<float>& th) {
  uint64_t h = 1469598103934665603ull;  // FNV-1a over the angle bytes
  const unsigned char* p = reinterpret_cast<const unsigned char*>(th.data());
  for (size_t i = 0; i < th.size() * sizeof(float); ++i) { h ^= p[i]; h *= 1099511628211ull; }
  return h;
}

// Average ms per slice of one radon of slice sl with variant v, as a batched
// launch would run it: the cells pass (when v reads it) once, plus the
// projector launched over R copies of the slice divided by R, R the number of
// the batch's slices of this shape, at least enough for two workgroups per
// slot of the GPU (a single small slice would leave CUs idle that a batch
// fills), at most 8.  Each copy reads an image of its own, of ones: the LDS
// projectors skip boxes of zeros (an FBP image is zero outside the
// reconstruction disk, a TV iterate soon is not), and copies of one image
// would share L2 as a batch's slices do not.  The copies write scratch
// outputs.
double time_variant(acu_batch* b, int sl, int v, int reps) {
  Slice& s = b->s[sl];
  acu_ctx* c = b->ctx;
  const int saved = s.fpv;
  s.fpv = v;
  const bool cells = v == 0 || fpl_may_overflow(s.N, s.T, s.theta.data(), v);
  DevArray<CellTask> cl;
  DevArray<FPLTask> fl;
  DevArray<FPTask> fp;
  DevArray<int> ord;
  struct Scratch { float* p = nullptr; ~Scratch() { if (p) hipFree(p); } } sc, ones;
  int nfl = 0, nfp = 0, nfs = 0, R = 1;
  if (v > 0) {
    const int per = fpl_blocks(s.N, s.T, v);
    const int slots = v == FPV_CELL ? 256 : 1024;  // workgroups resident on the GPU at once
    int same = 0;
    for (const auto& o : b->s) same += o.N == s.N && o.T == s.T;
    R = std::max(1, std::min(8, std::max(same, (2 * slots + per - 1) / per)));
  }
  const size_t NN = (size_t)s.N * s.N;
  ACU_HIP_CHECK(hipMalloc(&ones.p, (size_t)R * NN * sizeof(float)));
  ACU_HIP_CHECK(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(ones.p), 0x3f800000u, (size_t)R * NN,
                                  c->stream));
  CellTask ct = cell_task(s, ones.p);
  ct.block_begin = 0;
  cl.upload({ct}, b->ctx->stream);
  if (v > 0) {
    const int per = fpl_blocks(s.N, s.T, v);
    if (R > 1) ACU_HIP_CHECK(hipMalloc(&sc.p, (size_t)(R - 1) * s.N * s.T * sizeof(float)));
    std::vector<FPLTask> tv;
    for (int r = 0; r < R; ++r) {
      FPLTask t = fpl_task(s, ones.p + (size_t)r * NN, r == 0 ? s.ruk : sc.p + (size_t)(r - 1) * s.N * s.T);
      t.block_begin = r * per;
      tv.push_back(t);
    }
    fl.upload(tv, b->ctx->stream);
    nfl = R * per;
    if (b->fpl_order && v != FPV_CELL) {
      std::vector<FplGeom> geo;
      for (const FPLTask& t : tv) geo.push_back({t.N, t.nab, t.block_begin});
      ord.upload(fpl_dispatch_order(geo, 0, nfl), b->ctx->stream);
    }
  } else {
    FPTask t = fp_task(s, s.ruk);
    t.block_begin = 0;
    t.sum_begin = 0;
    fp.upload({t}, b->ctx->stream);
    nfp = fp_ray_blocks(s.N) * s.T * fp_chunks(s.N, s.fp_clen);
    nfs = fp_sum_blocks(s.N, s.T);
  }
  auto timed_ms = [&](auto&& launch) {
    hipEvent_t e0 = c->ev(), e1 = c->ev();
    launch();  // warm-up
    ACU_HIP_CHECK(hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps; ++r) launch();
    ACU_HIP_CHECK(hipEventRecord(e1, c->stream));
    check_launch();
    ACU_HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    ACU_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    c->pool.push_back(e0);
    c->pool.push_back(e1);
    return (double)ms / reps;
  };
  double ms = 0.0;
  if (cells) ms += timed_ms([&] { launch_cells(cl.d, 1, 0, cell_blocks(s.N), c->stream); });
  if (v > 0) ms += timed_ms([&] { launch_forward_lds(fl.d, R, 0, nfl, v, ord.d, c->stream); }) / R;
  else ms += timed_ms([&] { launch_forward(fp.d, 1, 0, nfp, 0, nfs, c->stream); });
  cl.release(); fl.release(); fp.release(); ord.release();
  s.fpv = saved;
  return ms;
}

const char* variant_name(int v) {
  return v == 0 ? "gather" : v == 1 ? "lds1" : v == 2 ? "lds2" : v == FPV_CELL ? "cell16" : "lds4";
}

const char* ramp_name(int v) {
  return v == RAMP_FP64 ? "fp64" : v == RAMP_FP32 ? "fp32" : v == RAMP_BF16X3 ? "bf16x3"
         : v == RAMP_BF16X6 ? "bf16x6" : "fft";
}

// Average ms of one ramp filter of every slice (f -> sinoCS) with variant v.
double time_ramp(acu_batch* b, int v, int reps) {
  acu_ctx* c = b->ctx;
  const int saved = b->ramp_variant;
  b->ramp_variant = v;
  auto run = [&] {
    if (v == RAMP_FFT) launch_ramp_fft(b->rf_f.ff.d, b->rf_f.ff.n, 0, b->ff_total, b->ff_maxL, c->stream);
    else launch_ramp(b->rf_f.rf.d, b->rf_f.rf.n, 0, b->rf_total, v, c->stream);
  };
  hipEvent_t e0 = c->ev(), e1 = c->ev();
  run();
  ACU_HIP_CHECK(hipEventRecord(e0, c->stream));
  for (int r = 0; r < reps; ++r) run();
  ACU_HIP_CHECK(hipEventRecord(e1, c->stream));
  check_launch();
  ACU_HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  ACU_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  c->pool.push_back(e0);
  c->pool.push_back(e1);
  b->ramp_variant = saved;
  return (double)ms / reps;
}

// Slices the backprojection timing covers: the batch's first slices up to
// 8192 tiles of 64 x 32 (about five rounds of the GPU's workgroup slots),
// at least one -- enough to time a full GPU without launching the whole of
// a large batch three times.
int bp_timing_slices(const acu_batch* b) {
  int n = 0, tiles = 0;
  while (n < (int)b->s.size() && (n == 0 || tiles < 8192)) {
    tiles += b->s[n].ntx * bp_tiles_y(b->s[n].N, 32);
    ++n;
  }
  return n;
}

// Average ms of one batched backprojection of those slices (sinoCS ->
// scratch images, contents irrelevant to the timing) at tile height ty.
double time_bp(acu_batch* b, int ty, int reps) {
  acu_ctx* c = b->ctx;
  const int ns = bp_timing_slices(b);
  size_t total = 0;
  for (int i = 0; i < ns; ++i) total += (size_t)b->s[i].N * b->s[i].N;
  struct Scratch { float* p = nullptr; ~Scratch() { if (p) hipFree(p); } } sc;
  ACU_HIP_CHECK(hipMalloc(&sc.p, total * sizeof(float)));
  std::vector<BPTask> tv;
  int tiles = 0;
  size_t off = 0;
  for (int i = 0; i < ns; ++i) {
    const Slice& s = b->s[i];
    BPTask t = bp_task(s, s.sinoCS, sc.p + off);
    t.nty = bp_tiles_y(s.N, ty);
    t.tile_begin = tiles;
    tiles += s.ntx * t.nty;
    off += (size_t)s.N * s.N;
    tv.push_back(t);
  }
  DevArray<BPTask> arr;
  arr.upload(tv, b->ctx->stream);
  std::vector<float> w;
  for (int i = 0; i < ns; ++i) bp_tile_work(b->s[i], ty, w);
  const BPShares sh = bp_xcd_shares(w);
  auto run = [&] { launch_backproject(arr.d, arr.n, 0, tiles, ty, b->bp_refpos, sh, c->stream); };
  hipEvent_t e0 = c->ev(), e1 = c->ev();
  run();
  ACU_HIP_CHECK(hipEventRecord(e0, c->stream));
  for (int r = 0; r < reps; ++r) run();
  ACU_HIP_CHECK(hipEventRecord(e1, c->stream));
  check_launch();
  ACU_HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  ACU_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  c->pool.push_back(e0);
  c->pool.push_back(e1);
  arr.release();
  return (double)ms / reps;
}

// What autotune times: the projector variants (per slice shape), the ramp
// arithmetic and the backprojection tile height (per batch).
constexpr unsigned TUNE_FP = 1, TUNE_RAMP = 2, TUNE_BP = 4, TUNE_ALL = 7;

std::string autotune(acu_batch* b, int reps, unsigned what = TUNE_ALL) {
  require_theta(b);
  b->ctx->sync();
  static const int cand[5] = {1, 2, 4, FPV_CELL, 0};
  constexpr int NC = 5;
  std::map<std::tuple<int, int, uint64_t>, std::pair<int, std::vector<double>>> seen;
  std::ostringstream js;
  js << "{\"reps\":" << reps << ",\"slices\":[";
  for (size_t i = 0; i < ((what & TUNE_FP) ? b->s.size() : 0); ++i) {
    Slice& s = b->s[i];
    const auto key = std::make_tuple(s.N, s.T, theta_hash(s.theta));
    auto it = seen.find(key);
    if (it == seen.end()) {
      std::vector<double> ms(NC);
      int best = 0;
      for (int k = 0; k < NC; ++k) {
        ms[k] = time_variant(b, (int)i, cand[k], reps);
        if (ms[k] < ms[best]) best = k;
      }
      it = seen.emplace(key, std::make_pair(cand[best], ms)).first;
    }
    s.fpv = it->second.first;
    js << (i ? "," : "") << "{\"slice\":" << i << ",\"N\":" << s.N << ",\"T\":" << s.T << ",\"ms\":{";
    for (int k = 0; k < NC; ++k)
      js << (k ? "," : "") << "\"" << variant_name(cand[k]) << "\":" << it->second.second[k];
    js << "},\"choice\":\"" << variant_name(s.fpv) << "\"}";
  }
  js << "]";
  // ramp filter: the fastest arithmetic among those that meet the FBP / TV
  // accuracy bar (bf16x3, with ~16 bits per operand, does not)
  if (what & TUNE_RAMP) {
    static const int rc[4] = {RAMP_FP64, RAMP_FP32, RAMP_BF16X6, RAMP_FFT};
    bool fft_ok = true;
    for (auto& s : b->s) fft_ok = fft_ok && s.fftG != nullptr;
    int best = -1;
    double best_ms = 0.0;
    js << ",\"ramp\":{\"ms\":{";
    bool first = true;
    for (int k = 0; k < 4; ++k) {
      if (rc[k] == RAMP_FFT && !fft_ok) continue;
      const double ms = time_ramp(b, rc[k], reps);
      js << (first ? "" : ",") << "\"" << ramp_name(rc[k]) << "\":" << ms;
      first = false;
      if (best < 0 || ms < best_ms) { best = rc[k]; best_ms = ms; }
    }
    b->ramp_variant = best;
    js << "},\"choice\":\"" << ramp_name(best) << "\"}";
  }
  // backprojection tile height: the fastest over the batch's first slices
  if (what & TUNE_BP) {
    static const int tys[3] = {16, 32, 64};
    int best = b->bp_ty;
    double best_ms = -1.0;
    js << ",\"bp\":{\"ms\":{";
    for (int k = 0; k < 3; ++k) {
      const double ms = time_bp(b, tys[k], reps);
      js << (k ? "," : "") << "\"" << tys[k] << "\":" << ms;
      if (best_ms < 0.0 || ms < best_ms) { best = tys[k]; best_ms = ms; }
    }
    js << "},\"choice\":" << best << "}";
    if (best != b->bp_ty) set_bp_tile(b, best);
  }
  js << "}";
  refresh_fp(b);
  return js.str();
}

// ---- differentiation: exact adjoint and the gradients of
// L(x, s) = 1/2 |shift_s(radon(x)) - b|^2 + lambda TV_eps(x), b = the slice's f ----
AdjTask adj_task(const Slice& s, int block_begin, const float* sino, float* out) {
  return AdjTask{s.N, s.T, s.mid, block_begin, adj_tiles(s.N), sino, s.scfp, out};
}

void radon_adjoint_dev(acu_batch* b, int sl, float* img, const float* sino) {
  Slice& s = slice_of(b, sl);
  acu_ctx* c = b->ctx;
  AdjTask* d = put_adhoc(b, adj_task(s, 0, sino, img));
  const int nb = adj_tiles(s.N) * adj_tiles(s.N);
  timed(c, K_ADJ, [&] { launch_adjoint(d, 1, 0, nb, c->stream); });
  check_launch();
}

void layout_grad(acu_batch* b, Carver& cv) {
  for (auto& s : b->s) {
    const size_t NN = (size_t)s.N * s.N, NT = (size_t)s.N * s.T;
    s.gx_x = cv.take<float>(NN);
    s.gx_g = cv.take<float>(NN);
    s.gx_y = cv.take<float>(NT);
    s.gx_rs = cv.take<float>(NT);
    s.gx_sh = cv.take<float>((size_t)s.T);
    s.gx_gs = cv.take<float>((size_t)s.T);
    s.gx_hs = cv.take<float>((size_t)s.T);
    s.gx_pd = cv.take<double>((size_t)s.T);
    s.gx_pt = cv.take<double>((size_t)tvg_blocks(s.N));
    s.gx_loss = cv.take<double>(4);
  }
}

void ensure_grad(acu_batch* b) {
  if (b->grad_mem) return;
  Carver dry;
  layout_grad(b, dry);
  ACU_HIP_CHECK(hipMalloc(&b->grad_mem, dry.off + 256));
  Carver cv{static_cast<char*>(b->grad_mem), 0};
  layout_grad(b, cv);
}

// radon of every slice: imgs[i] -> sinos[i] (the batch's block ranges)
void radon_many(acu_batch* b, const std::vector<const float*>& imgs, const std::vector<float*>& sinos) {
  std::vector<const float*> im(imgs);
  std::vector<float*> out(sinos);
  std::map<const Slice*, int> idx;
  for (size_t i = 0; i < b->s.size(); ++i) idx[&b->s[i]] = (int)i;
  build_fpset(b, b->fp_g, [&](Slice& s) { return im[idx[&s]]; }, [&](Slice& s) { return out[idx[&s]]; });
  radon_seq(b, b->fp_g, -1);
}

// x[i], shift[i] (nullable), gx[i], gs[i], hs[i] (nullable) are device
// pointers; gx may be empty (no image gradient).  loss receives 3 doubles
// per slice.
void loss_grad_run(acu_batch* b, const std::vector<const float*>& x, const std::vector<const float*>& shift,
                   float lambda, float eps, const std::vector<float*>& gx, const std::vector<float*>& gs,
                   const std::vector<float*>& hs, double* loss) {
  acu_ctx* c = b->ctx;
  require_theta(b);
  ensure_grad(b);
  const int n = (int)b->s.size();
  std::vector<float*> ys(n);
  for (int i = 0; i < n; ++i) ys[i] = b->s[i].gx_y;
  radon_many(b, x, ys);
  std::vector<ResTask> res(n);
  std::vector<TVGTask> tvg(n);
  std::vector<LossTask> lt(n);
  std::vector<AdjTask> adj;
  int rows = 0, tvb = 0, adjb = 0;
  for (int i = 0; i < n; ++i) {
    Slice& s = b->s[i];
    res[i] = ResTask{s.N, s.T, rows, s.gx_y, s.f, shift[i], s.gx_rs, gs[i], hs[i], s.gx_pd};
    rows += s.T;
    const int nb = tvg_blocks(s.N);
    tvg[i] = TVGTask{s.N, tvb, nb, lambda, eps, x[i], gx.empty() ? nullptr : gx[i], s.gx_pt};
    tvb += nb;
    lt[i] = LossTask{s.T, nb, (double)lambda, s.gx_pd, s.gx_pt, s.gx_loss};
    if (!gx.empty()) {
      adj.push_back(adj_task(s, adjb, s.gx_rs, gx[i]));
      adjb += adj_tiles(s.N) * adj_tiles(s.N);
    }
  }
  b->res_g.upload(res, b->ctx->stream);
  b->tvg_g.upload(tvg, b->ctx->stream);
  b->loss_g.upload(lt, b->ctx->stream);
  launch_shift_resid(b->res_g.d, n, 0, rows, c->stream);
  check_launch();
  if (!adj.empty()) {
    b->adj_g.upload(adj, b->ctx->stream);
    timed(c, K_ADJ, [&] { launch_adjoint(b->adj_g.d, n, 0, adjb, c->stream); });
    check_launch();
  }
  launch_tvgrad(b->tvg_g.d, n, 0, tvb, c->stream);
  launch_loss_finalize(b->loss_g.d, n, c->stream);
  check_launch();
  for (int i = 0; i < n; ++i)
    ACU_HIP_CHECK(hipMemcpyAsync(loss + 3 * i, b->s[i].gx_loss, 3 * sizeof(double), hipMemcpyDeviceToHost,
                                 c->stream));
  c->sync();
}

}  // namespace
}  // namespace acu

using acu::Error;

#define ACU_TRY(...)                                                         \
  try {                                                                      \
    __VA_ARGS__;                                                             \
    return ACU_OK;                                                           \
  } catch (const Error& e) {                                                 \
    return fail(ACU_ERR_HIP, e.msg);                                         \
  } catch (const std::bad_alloc&) {                                          \
    return fail(ACU_ERR_HIP, "host allocation failed");                      \
  }

namespace acu {
int ctx_device(const acu_ctx* c) { return c->device; }
hipStream_t ctx_stream(const acu_ctx* c) { return c->stream; }
int record_error(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace acu

extern "C" {

const char* acu_last_error(void) { return g_err.c_str(); }

int acu_ctx_create(int device, acu_ctx** out) {
  if (!out) return fail(ACU_ERR_ARG, "acu_ctx_create: out is NULL");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(ACU_ERR_HIP, "no HIP device available");
  if (device < 0 || device >= n) return fail(ACU_ERR_ARG, "acu_ctx_create: device index out of range");
  auto* c = new acu_ctx();
  c->device = device;
  try {
    ACU_HIP_CHECK(hipSetDevice(device));
    ACU_HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    acu::noalloc_setup(device);  // the device address of this GPU's no-alloc scratch
  } catch (const Error& e) {
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return fail(ACU_ERR_HIP, e.msg);
  }
  *out = c;
  return ACU_OK;
}

int acu_ctx_destroy(acu_ctx* ctx) {
  if (!ctx) return ACU_OK;
  hipStreamSynchronize(ctx->stream);
  for (auto& p : ctx->pending) { hipEventDestroy(p.a); hipEventDestroy(p.b); }
  for (auto e : ctx->pool) hipEventDestroy(e);
  for (auto e : ctx->pin_ev)
    if (e) hipEventDestroy(e);
  for (auto b : ctx->pin)
    if (b) hipHostFree(b);
  hipStreamDestroy(ctx->stream);
  delete ctx;
  return ACU_OK;
}

int acu_ctx_synchronize(acu_ctx* ctx) {
  if (!ctx) return fail(ACU_ERR_ARG, "null context");
  ACU_TRY(hipSetDevice(ctx->device); ctx->sync());
}

int acu_ctx_set_timing(acu_ctx* ctx, int enable) {
  if (!ctx) return fail(ACU_ERR_ARG, "null context");
  ACU_TRY(ctx->sync(); ctx->timing = enable != 0;
          for (int k = 0; k < acu_ctx::KINDS; ++k) { ctx->ms[k] = 0; ctx->count[k] = 0; });
}

int acu_ctx_kernel_time(acu_ctx* ctx, int kind, double* total_ms, long long* launches) {
  if (!ctx || kind < 0 || kind >= acu_ctx::KINDS) return fail(ACU_ERR_ARG, "acu_ctx_kernel_time: bad argument");
  ACU_TRY(ctx->sync(); if (total_ms) *total_ms = ctx->ms[kind];
          if (launches) *launches = ctx->count[kind]);
}

int acu_batch_create(acu_ctx* ctx, int nslices, const int* nimage, const int* ntheta,
                     const int* wpu, const float* const* theta_deg, acu_batch** out) {
  if (!ctx || !out || nslices <= 0 || !nimage || !ntheta || !wpu)
    return fail(ACU_ERR_ARG, "acu_batch_create: bad argument");
  *out = nullptr;
  for (int i = 0; i < nslices; ++i) {
    if (nimage[i] < 3 || nimage[i] > 16000) return fail(ACU_ERR_ARG, "nimage must be in [3, 16000]");
    if (ntheta[i] < 1 || ntheta[i] > 1 << 20) return fail(ACU_ERR_ARG, "ntheta must be in [1, 2^20]");
    if (wpu[i] < 1) return fail(ACU_ERR_ARG, "wpu must be >= 1 (the mask divides by it)");
  }
  std::unique_ptr<acu_batch> b(new acu_batch());
  b->ctx = ctx;
  try {
    ACU_HIP_CHECK(hipSetDevice(ctx->device));
    b->s.resize(nslices);
    int rf = 0, ew = 0, sw = 0;
    // forward-projection variant per slice: ACUMEN_FP_PATH = lds (default) |
    // gather; ACUMEN_FPL_AL = 1 | 2 | 4 forces the LDS angles per lane;
    // acu_batch_autotune replaces both by measurement
    const char* pv = std::getenv("ACUMEN_FP_PATH");
    const bool gather = pv && std::strcmp(pv, "gather") == 0;
    const char* av = std::getenv("ACUMEN_FPL_AL");
    const int al_env = av ? std::atoi(av) : 0;
    // ramp-filter arithmetic: ACUMEN_RAMP = fp64 (default) | fp32 | bf16x3 | bf16x6 | fft
    if (const char* rv = std::getenv("ACUMEN_RAMP")) {
      const std::string r(rv);
      b->ramp_variant = r == "fp32" ? acu::RAMP_FP32 : r == "bf16x3" ? acu::RAMP_BF16X3
                        : r == "bf16x6" ? acu::RAMP_BF16X6 : r == "fft" ? acu::RAMP_FFT : acu::RAMP_FP64;
    }
    // backprojection positions: ACUMEN_BP_POSITIONS = reference (default) |
    // tile for FBP images, ACUMEN_TV_BP_POSITIONS = reference (default) |
    // tile inside tvopt's iteration
    if (const char* pv2 = std::getenv("ACUMEN_BP_POSITIONS")) b->bp_refpos = std::strcmp(pv2, "tile") != 0;
    if (const char* pv3 = std::getenv("ACUMEN_TV_BP_POSITIONS")) b->tv_bp_refpos = std::strcmp(pv3, "tile") != 0;
    // LDS projector dispatch: ACUMEN_FPL_ORDER = balanced (default:
    // fpl_dispatch_order) | rb (ray blocks in order)
    if (const char* po = std::getenv("ACUMEN_FPL_ORDER")) b->fpl_order = std::strcmp(po, "rb") != 0;
    // backprojection tile height: ACUMEN_BP_TILE = 16 | 32 (default) | 64
    if (const char* tv = std::getenv("ACUMEN_BP_TILE")) {
      const int ty = std::atoi(tv);
      if (acu::bp_tile_ok(ty)) b->bp_ty = ty;
    }
    // operator schedule of tvopt: ACUMEN_TV_SCHEDULE = linear (default) |
    // reuse | reference; ACUMEN_TV_REFRESH = exact refresh period of the
    // linear schedule in outer iterations (default 0: none)
    if (const char* sv = std::getenv("ACUMEN_TV_SCHEDULE")) {
      const std::string v(sv);
      b->tv_sched = v == "reference" ? ACU_TV_SCHED_REFERENCE : v == "reuse" ? ACU_TV_SCHED_REUSE
                                                                             : ACU_TV_SCHED_LINEAR;
    }
    if (const char* rv = std::getenv("ACUMEN_TV_REFRESH")) b->tv_refresh = std::max(0, std::atoi(rv));
    // tuning: forward-projection sweep chunk length (steps), ACUMEN_FP_CHUNK
    const char* ev = std::getenv("ACUMEN_FP_CHUNK");
    const int fp_clen_env = ev ? std::atoi(ev) : 0;
    for (int i = 0; i < nslices; ++i) {
      auto& s = b->s[i];
      s.N = nimage[i]; s.T = ntheta[i]; s.wpu = wpu[i];
      s.mid = (float)((s.N + 1) / 2);
      s.ntx = acu::bp_tiles_x(s.N); s.nsub = (s.N + acu::BP_SUB - 1) / acu::BP_SUB;
      s.theta.assign(s.T, 0.0f);
      s.rf_begin = rf; s.rf_n = acu::rf_tiles_m(s.N) * acu::rf_tiles_n(s.T); rf += s.rf_n;
      s.ff_begin = b->ff_total; s.ff_n = (s.T + 1) / 2; b->ff_total += s.ff_n;
      b->ff_maxL = std::max(b->ff_maxL, acu::fft_len(s.N));
      s.fp_clen = fp_clen_env > 0 ? fp_clen_env : acu::FP_CHUNK_DEFAULT;
      s.fpv = gather ? 0
                     : (al_env > 0 && acu::fpv_valid(al_env)) ? al_env
                                                                   : acu::fpl_angles_per_lane(s.N, s.T);
      s.ew_begin = ew; s.ew_n = acu::ew_blocks((size_t)s.N * s.N); ew += s.ew_n;
      s.sw_begin = sw; s.sw_n = acu::ew_blocks((size_t)s.N * s.T); sw += s.sw_n;
    }
    b->rf_total = rf;
    acu::assign_bp(b.get());
    b->ew_total = ew; b->sw_total = sw;
    acu::Carver dry;
    acu::layout_ops(b.get(), dry);
    ACU_HIP_CHECK(hipMalloc(&b->ops_mem, dry.off + 256));
    acu::Carver cv{static_cast<char*>(b->ops_mem), 0};
    acu::layout_ops(b.get(), cv);
    std::map<std::tuple<int, int, int>, int> built;  // (N, T, wpu) -> first slice with it
    for (int i = 0; i < nslices; ++i) {
      const auto key = std::make_tuple(nimage[i], ntheta[i], wpu[i]);
      const auto it = built.find(key);
      if (it == built.end()) {
        acu::upload_tables(b.get(), i);
        built.emplace(key, i);
      } else {
        acu::copy_tables(b.get(), it->second, i);
      }
      ACU_HIP_CHECK(hipMemsetAsync(b->s[i].f, 0, (size_t)b->s[i].N * b->s[i].T * 4, b->ctx->stream));
      if (theta_deg && theta_deg[i]) {
        std::memcpy(b->s[i].theta.data(), theta_deg[i], sizeof(float) * b->s[i].T);
        acu::upload_theta(b.get(), i);
      }
    }
    acu::assign_bp_shares(b.get());
    if (b->ramp_variant == acu::RAMP_FFT)
      for (auto& s : b->s)
        if (!s.fftG)
          throw Error("ACUMEN_RAMP=fft: the FFT filter supports N <= " + std::to_string(acu::FFT_MAX_L / 2 + 1));
    acu::assign_fp(b.get());
    acu::build_op_tasks(b.get());
  } catch (const Error& e) {
    return fail(ACU_ERR_HIP, e.msg);
  }
  *out = b.release();
  return ACU_OK;
}

int acu_plan_create(acu_ctx* ctx, int nimage, int ntheta, int wpu, const float* theta_deg,
                    acu_batch** out) {
  const float* th[1] = {theta_deg};
  return acu_batch_create(ctx, 1, &nimage, &ntheta, &wpu, th, out);
}

int acu_batch_destroy(acu_batch* b) {
  if (!b) return ACU_OK;
  hipSetDevice(b->ctx->device);
  hipStreamSynchronize(b->ctx->stream);
  delete b;
  return ACU_OK;
}

int acu_batch_nslices(const acu_batch* b) { return b ? (int)b->s.size() : 0; }

int acu_batch_set_theta(acu_batch* b, int slice, const float* theta_deg) {
  if (!b || !theta_deg || slice < 0 || slice >= (int)b->s.size())
    return fail(ACU_ERR_ARG, "acu_batch_set_theta: bad argument");
  // results of the old angles no longer describe this slice
  b->s[slice].fbp_valid = false;
  b->s[slice].tv_valid = false;
  ACU_TRY(hipSetDevice(b->ctx->device); b->ctx->sync();
          std::memcpy(b->s[slice].theta.data(), theta_deg, sizeof(float) * b->s[slice].T);
          acu::upload_theta(b, slice);
          // whether the cells pass is launched depends on the angles
          acu::refresh_fp(b));
}

int acu_batch_set_sino(acu_batch* b, int slice, const float* f_host) {
  if (!b || !f_host || slice < 0 || slice >= (int)b->s.size())
    return fail(ACU_ERR_ARG, "acu_batch_set_sino: bad argument");
  auto& s = b->s[slice];
  // results of the old sinogram no longer describe this slice
  s.fbp_valid = false;
  s.tv_valid = false;
  ACU_TRY(hipSetDevice(b->ctx->device);
          ACU_HIP_CHECK(hipMemcpyAsync(s.f, f_host, (size_t)s.N * s.T * 4, hipMemcpyHostToDevice,
                                       b->ctx->stream));
          b->ctx->sync());
}

int acu_batch_get_image(acu_batch* b, int slice, int which, float* img_host) {
  if (!b || !img_host || slice < 0 || slice >= (int)b->s.size())
    return fail(ACU_ERR_ARG, "acu_batch_get_image: bad argument");
  auto& s = b->s[slice];
  const float* src = nullptr;
  if (which == ACU_IMG_FBP) {
    if (!s.fbp_valid) return fail(ACU_ERR_STATE, "no FBP image computed yet");
    src = s.fbp;
  } else if (which == ACU_IMG_TV) {
    if (!s.tv_valid) return fail(ACU_ERR_STATE, "no TV image computed yet");
    src = s.uk;
  } else {
    return fail(ACU_ERR_ARG, "acu_batch_get_image: which must be ACU_IMG_TV or ACU_IMG_FBP");
  }
  ACU_TRY(hipSetDevice(b->ctx->device);
          ACU_HIP_CHECK(hipMemcpyAsync(img_host, src, (size_t)s.N * s.N * 4, hipMemcpyDeviceToHost,
                                       b->ctx->stream));
          b->ctx->sync());
}

float* acu_batch_device_ptr(acu_batch* b, int slice, int which) {
  if (!b || slice < 0 || slice >= (int)b->s.size()) { fail(ACU_ERR_ARG, "bad slice"); return nullptr; }
  auto& s = b->s[slice];
  if (which == ACU_SINO_F) return s.f;
  if (which == ACU_IMG_FBP) return s.fbp;
  if (which == ACU_IMG_TV) {
    try { acu::ensure_tv(b, b->tv_delta_cap > 0 ? b->tv_delta_cap : 1); }
    catch (const Error& e) { fail(ACU_ERR_HIP, e.msg); return nullptr; }
    return s.uk;
  }
  fail(ACU_ERR_ARG, "bad selector");
  return nullptr;
}

int acu_batch_fbp(acu_batch* b) {
  if (!b) return fail(ACU_ERR_ARG, "null batch");
  ACU_TRY(hipSetDevice(b->ctx->device); acu::require_theta(b); acu::fbp_all(b, -1);
          for (auto& s : b->s) s.fbp_valid = true);
}

int acu_batch_tvopt(acu_batch* b, float mu, float lambda, int nouter, int ninner, int ncg, int mode) {
  if (!b) return fail(ACU_ERR_ARG, "null batch");
  if (mode != ACU_TV_PLAIN && mode != ACU_TV_BATCHED && mode != ACU_TV_GRAPH)
    return fail(ACU_ERR_ARG, "unknown tvopt mode");
  if (lambda == 0.0f) return fail(ACU_ERR_ARG, "lambda must be nonzero (the shrink threshold is 1/lambda)");
  ACU_TRY(hipSetDevice(b->ctx->device); acu::tvopt_run(b, mu, lambda, nouter, ninner, ncg, mode));
}

int acu_batch_set_tv_schedule(acu_batch* b, int sched, int refresh) {
  if (!b) return fail(ACU_ERR_ARG, "null batch");
  if (sched != ACU_TV_SCHED_REFERENCE && sched != ACU_TV_SCHED_REUSE && sched != ACU_TV_SCHED_LINEAR)
    return fail(ACU_ERR_ARG, "acu_batch_set_tv_schedule: unknown schedule");
  if (refresh < 0) return fail(ACU_ERR_ARG, "acu_batch_set_tv_schedule: refresh must be >= 0");
  ACU_TRY(hipSetDevice(b->ctx->device); b->ctx->sync(); b->tv_sched = sched; b->tv_refresh = refresh;
          b->drop_graphs(););
}

int acu_batch_set_bp_positions(acu_batch* b, int fbp, int tv) {
  if (!b) return fail(ACU_ERR_ARG, "null batch");
  auto ok = [](int v) { return v == ACU_BP_POS_REFERENCE || v == ACU_BP_POS_TILE; };
  if (!ok(fbp) || !ok(tv)) return fail(ACU_ERR_ARG, "acu_batch_set_bp_positions: unknown positions");
  // the TV task table points at s.fbp or s.ftv: tvopt rebuilds it
  ACU_TRY(hipSetDevice(b->ctx->device); b->ctx->sync(); b->bp_refpos = fbp == ACU_BP_POS_REFERENCE;
          b->tv_bp_refpos = tv == ACU_BP_POS_REFERENCE; b->tv.release(); b->drop_graphs();
          for (auto& s : b->s) { s.fbp_valid = false; s.tv_valid = false; });
}

int acu_batch_bp_positions(const acu_batch* b, int* fbp, int* tv) {
  if (!b) return fail(ACU_ERR_ARG, "null batch");
  if (fbp) *fbp = b->bp_refpos ? ACU_BP_POS_REFERENCE : ACU_BP_POS_TILE;
  if (tv) *tv = b->tv_bp_refpos ? ACU_BP_POS_REFERENCE : ACU_BP_POS_TILE;
  return ACU_OK;
}

int acu_batch_tv_schedule(const acu_batch* b, int* refresh) {
  if (!b) return -fail(ACU_ERR_ARG, "null batch");
  if (refresh) *refresh = b->tv_refresh;
  return b->tv_sched;
}

int acu_batch_get_delta(acu_batch* b, int slice, float* delta_host, int n) {
  if (!b || !delta_host || slice < 0 || slice >= (int)b->s.size() || n < 0)
    return fail(ACU_ERR_ARG, "acu_batch_get_delta: bad argument");
  auto& s = b->s[slice];
  if (!s.delta || !s.tv_valid || b->last_nouter < 0)
    return fail(ACU_ERR_STATE, "no tvopt has completed on this slice since its sinogram or angles were set");
  if (n > b->last_nouter)
    return fail(ACU_ERR_ARG, "acu_batch_get_delta: n exceeds the last tvopt's nouter (" +
                                 std::to_string(b->last_nouter) + ")");
  ACU_TRY(hipSetDevice(b->ctx->device);
          ACU_HIP_CHECK(hipMemcpyAsync(delta_host, s.delta, (size_t)n * 4, hipMemcpyDeviceToHost,
                                       b->ctx->stream));
          b->ctx->sync());
}

int acu_radon_dev(acu_batch* b, int slice, float* sino_dev, const float* img_dev) {
  if (!b || !sino_dev || !img_dev) return fail(ACU_ERR_ARG, "acu_radon_dev: null argument");
  ACU_TRY(hipSetDevice(b->ctx->device); acu::radon_dev(b, slice, sino_dev, img_dev));
}

int acu_ramp_dev(acu_batch* b, int slice, float* out_dev, const float* sino_dev) {
  if (!b || !out_dev || !sino_dev) return fail(ACU_ERR_ARG, "acu_ramp_dev: null argument");
  ACU_TRY(hipSetDevice(b->ctx->device); acu::ramp_dev(b, slice, out_dev, sino_dev));
}

int acu_backproject_dev(acu_batch* b, int slice, float* img_dev, const float* filtered_dev) {
  if (!b || !img_dev || !filtered_dev) return fail(ACU_ERR_ARG, "acu_backproject_dev: null argument");
  ACU_TRY(hipSetDevice(b->ctx->device); acu::bp_dev(b, slice, img_dev, filtered_dev));
}

int acu_iradon_dev(acu_batch* b, int slice, float* img_dev, const float* sino_dev) {
  if (!b || !img_dev || !sino_dev) return fail(ACU_ERR_ARG, "acu_iradon_dev: null argument");
  ACU_TRY(hipSetDevice(b->ctx->device);
          acu::Slice& s = acu::slice_of(b, slice);
          acu::ramp_dev(b, slice, s.sinoCS, sino_dev);
          acu::bp_dev(b, slice, img_dev, s.sinoCS));
}

namespace {
// Host-buffer round trip through device scratch for the single-slice forms.
struct DevBuf {
  float* p = nullptr;
  explicit DevBuf(size_t n) { ACU_HIP_CHECK(hipMalloc(&p, n * 4)); }
  ~DevBuf() { if (p) hipFree(p); }
};
}  // namespace

int acu_radon(acu_batch* b, float* sino_host, const float* img_host) {
  if (!b || !sino_host || !img_host) return fail(ACU_ERR_ARG, "acu_radon: null argument");
  ACU_TRY(hipSetDevice(b->ctx->device);
          acu::Slice& s = acu::slice_of(b, 0);
          DevBuf img((size_t)s.N * s.N), sino((size_t)s.N * s.T);
          ACU_HIP_CHECK(hipMemcpyAsync(img.p, img_host, (size_t)s.N * s.N * 4, hipMemcpyHostToDevice,
                                       b->ctx->stream));
          acu::radon_dev(b, 0, sino.p, img.p);
          ACU_HIP_CHECK(hipMemcpyAsync(sino_host, sino.p, (size_t)s.N * s.T * 4, hipMemcpyDeviceToHost,
                                       b->ctx->stream));
          b->ctx->sync());
}

int acu_iradon(acu_batch* b, float* img_host, const float* sino_host) {
  if (!b || !img_host || !sino_host) return fail(ACU_ERR_ARG, "acu_iradon: null argument");
  ACU_TRY(hipSetDevice(b->ctx->device);
          acu::Slice& s = acu::slice_of(b, 0);
          DevBuf img((size_t)s.N * s.N), sino((size_t)s.N * s.T);
          ACU_HIP_CHECK(hipMemcpyAsync(sino.p, sino_host, (size_t)s.N * s.T * 4, hipMemcpyHostToDevice,
                                       b->ctx->stream));
          acu::ramp_dev(b, 0, s.sinoCS, sino.p);
          acu::bp_dev(b, 0, img.p, s.sinoCS);
          ACU_HIP_CHECK(hipMemcpyAsync(img_host, img.p, (size_t)s.N * s.N * 4, hipMemcpyDeviceToHost,
                                       b->ctx->stream));
          b->ctx->sync());
}

int acu_get_slice(acu_batch* b, const float* f_host, float* img_fbp_host) {
  int rc = acu_batch_set_sino(b, 0, f_host);
  if (rc) return rc;
  if ((rc = acu_batch_fbp(b))) return rc;
  return acu_batch_get_image(b, 0, ACU_IMG_FBP, img_fbp_host);
}

int acu_batch_autotune(acu_batch* b, int reps, char* report, int cap) {
  if (!b || reps < 1 || (report && cap < 1)) return fail(ACU_ERR_ARG, "acu_batch_autotune: bad argument");
  ACU_TRY(hipSetDevice(b->ctx->device);
          const std::string js = acu::autotune(b, reps);
          if (report) {
            const size_t n = std::min(js.size(), (size_t)cap - 1);
            std::memcpy(report, js.data(), n);
            report[n] = '\0';
          });
}

int acu_batch_set_ramp_variant(acu_batch* b, int variant) {
  if (!b) return fail(ACU_ERR_ARG, "null batch");
  if (variant < acu::RAMP_FP64 || variant > acu::RAMP_FFT)
    return fail(ACU_ERR_ARG, "acu_batch_set_ramp_variant: unknown variant");
  if (variant == acu::RAMP_FFT)
    for (auto& s : b->s)
      if (!s.fftG)
        return fail(ACU_ERR_ARG, "acu_batch_set_ramp_variant: the FFT filter supports N <= " +
                                     std::to_string(acu::FFT_MAX_L / 2 + 1));
  ACU_TRY(hipSetDevice(b->ctx->device); b->ctx->sync(); b->ramp_variant = variant;
          b->drop_graphs(););
}

int acu_batch_ramp_variant(const acu_batch* b) {
  if (!b) return -fail(ACU_ERR_ARG, "null batch");
  return b->ramp_variant;
}

int acu_batch_set_bp_tile(acu_batch* b, int rows) {
  if (!b) return fail(ACU_ERR_ARG, "null batch");
  if (!acu::bp_tile_ok(rows)) return fail(ACU_ERR_ARG, "acu_batch_set_bp_tile: rows must be 16, 32 or 64");
  ACU_TRY(hipSetDevice(b->ctx->device); b->ctx->sync(); acu::set_bp_tile(b, rows));
}

int acu_batch_bp_tile(const acu_batch* b) {
  if (!b) return -fail(ACU_ERR_ARG, "null batch");
  return b->bp_ty;
}

int acu_batch_fp_variant(const acu_batch* b, int slice) {
  if (!b || slice < 0 || slice >= (int)b->s.size()) return -fail(ACU_ERR_ARG, "acu_batch_fp_variant: bad slice");
  return b->s[slice].fpv;
}

int acu_batch_set_fp_variant(acu_batch* b, int slice, int variant) {
  if (!b || slice < 0 || slice >= (int)b->s.size())
    return fail(ACU_ERR_ARG, "acu_batch_set_fp_variant: bad slice");
  if (!acu::fpv_valid(variant))
    return fail(ACU_ERR_ARG, "acu_batch_set_fp_variant: variant must be 0, 1, 2, 4 or 8");
  ACU_TRY(hipSetDevice(b->ctx->device); b->ctx->sync(); b->s[slice].fpv = variant; acu::refresh_fp(b));
}

int acu_radon_adjoint_dev(acu_batch* b, int slice, float* img_dev, const float* sino_dev) {
  if (!b || !img_dev || !sino_dev) return fail(ACU_ERR_ARG, "acu_radon_adjoint_dev: null argument");
  ACU_TRY(hipSetDevice(b->ctx->device); acu::radon_adjoint_dev(b, slice, img_dev, sino_dev));
}

int acu_radon_adjoint(acu_batch* b, float* img_host, const float* sino_host) {
  if (!b || !img_host || !sino_host) return fail(ACU_ERR_ARG, "acu_radon_adjoint: null argument");
  ACU_TRY(hipSetDevice(b->ctx->device);
          acu::Slice& s = acu::slice_of(b, 0);
          DevBuf img((size_t)s.N * s.N), sino((size_t)s.N * s.T);
          ACU_HIP_CHECK(hipMemcpyAsync(sino.p, sino_host, (size_t)s.N * s.T * 4, hipMemcpyHostToDevice,
                                       b->ctx->stream));
          acu::radon_adjoint_dev(b, 0, img.p, sino.p);
          ACU_HIP_CHECK(hipMemcpyAsync(img_host, img.p, (size_t)s.N * s.N * 4, hipMemcpyDeviceToHost,
                                       b->ctx->stream));
          b->ctx->sync());
}

namespace {
int check_grad_args(acu_batch* b, const float* const* x, float eps, double* loss, const char* who) {
  if (!b || !x || !loss) return fail(ACU_ERR_ARG, std::string(who) + ": null argument");
  if (!(eps > 0.0f)) return fail(ACU_ERR_ARG, std::string(who) + ": eps must be > 0");
  for (size_t i = 0; i < b->s.size(); ++i)
    if (!x[i]) return fail(ACU_ERR_ARG, std::string(who) + ": null image pointer");
  return ACU_OK;
}
}  // namespace

int acu_batch_loss_grad_dev(acu_batch* b, const float* const* x_dev, const float* const* shift_dev,
                            float lambda, float eps, float* const* grad_x_dev, float* const* grad_s_dev,
                            float* const* curv_s_dev, double* loss) {
  if (int rc = check_grad_args(b, x_dev, eps, loss, "acu_batch_loss_grad_dev")) return rc;
  const size_t n = b->s.size();
  if (grad_x_dev)
    for (size_t i = 0; i < n; ++i)
      if (!grad_x_dev[i]) return fail(ACU_ERR_ARG, "acu_batch_loss_grad_dev: null image-gradient pointer");
  ACU_TRY(hipSetDevice(b->ctx->device);
          std::vector<const float*> xs(x_dev, x_dev + n), sh(n, nullptr);
          std::vector<float*> gx, gs(n, nullptr), hs(n, nullptr);
          for (size_t i = 0; i < n; ++i) {
            if (shift_dev) sh[i] = shift_dev[i];
            if (grad_s_dev) gs[i] = grad_s_dev[i];
            if (curv_s_dev) hs[i] = curv_s_dev[i];
          }
          if (grad_x_dev) gx.assign(grad_x_dev, grad_x_dev + n);
          acu::loss_grad_run(b, xs, sh, lambda, eps, gx, gs, hs, loss));
}

int acu_batch_loss_grad(acu_batch* b, const float* const* x_host, const float* const* shift_host,
                        float lambda, float eps, float* const* grad_x_host, float* const* grad_s_host,
                        float* const* curv_s_host, double* loss) {
  if (int rc = check_grad_args(b, x_host, eps, loss, "acu_batch_loss_grad")) return rc;
  const size_t n = b->s.size();
  if (grad_x_host)
    for (size_t i = 0; i < n; ++i)
      if (!grad_x_host[i]) return fail(ACU_ERR_ARG, "acu_batch_loss_grad: null image-gradient pointer");
  ACU_TRY(hipSetDevice(b->ctx->device);
          acu::ensure_grad(b);
          acu_ctx* c = b->ctx;
          std::vector<const float*> xs(n), sh(n, nullptr);
          std::vector<float*> gx, gs(n, nullptr), hs(n, nullptr);
          for (size_t i = 0; i < n; ++i) {
            acu::Slice& s = b->s[i];
            ACU_HIP_CHECK(hipMemcpyAsync(s.gx_x, x_host[i], (size_t)s.N * s.N * 4, hipMemcpyHostToDevice,
                                         c->stream));
            xs[i] = s.gx_x;
            if (shift_host && shift_host[i]) {
              ACU_HIP_CHECK(hipMemcpyAsync(s.gx_sh, shift_host[i], (size_t)s.T * 4, hipMemcpyHostToDevice,
                                           c->stream));
              sh[i] = s.gx_sh;
            }
            if (grad_s_host && grad_s_host[i]) gs[i] = s.gx_gs;
            if (curv_s_host && curv_s_host[i]) hs[i] = s.gx_hs;
            if (grad_x_host) gx.push_back(s.gx_g);
          }
          c->sync();  // host buffers are not needed past this point
          acu::loss_grad_run(b, xs, sh, lambda, eps, gx, gs, hs, loss);
          for (size_t i = 0; i < n; ++i) {
            acu::Slice& s = b->s[i];
            if (grad_x_host)
              ACU_HIP_CHECK(hipMemcpyAsync(grad_x_host[i], s.gx_g, (size_t)s.N * s.N * 4,
                                           hipMemcpyDeviceToHost, c->stream));
            if (grad_s_host && grad_s_host[i])
              ACU_HIP_CHECK(hipMemcpyAsync(grad_s_host[i], s.gx_gs, (size_t)s.T * 4, hipMemcpyDeviceToHost,
                                           c->stream));
            if (curv_s_host && curv_s_host[i])
              ACU_HIP_CHECK(hipMemcpyAsync(curv_s_host[i], s.gx_hs, (size_t)s.T * 4, hipMemcpyDeviceToHost,
                                           c->stream));
          }
          c->sync());
}

namespace {
// One of the two lanes of the volume pipeline: its own context (stream),
// slice batches (full size and, lazily, the tail) and pinned staging.
struct Lane {
  acu_ctx* ctx = nullptr;
  acu_batch* full = nullptr;
  acu_batch* tail = nullptr;
  float *sin_pin = nullptr, *fbp_pin = nullptr, *tv_pin = nullptr;
  int group = -1, count = 0;
  ~Lane() {
    acu_batch_destroy(tail);
    acu_batch_destroy(full);
    acu_ctx_destroy(ctx);
    if (sin_pin) hipHostFree(sin_pin);
    if (fbp_pin) hipHostFree(fbp_pin);
    if (tv_pin) hipHostFree(tv_pin);
  }
};

void check_rc(int rc) {
  if (rc != ACU_OK) throw Error(acu_last_error());
}

// Host copy between caller memory and pinned staging, split over
// ACUMEN_COPY_THREADS threads (default: the host's hardware threads, at most
// 16): one thread's memcpy reaches only a fraction of the host memory
// bandwidth that the PCIe transfers need.
void par_copy(void* dst, const void* src, size_t bytes) {
  static const int nt = [] {
    const char* e = std::getenv("ACUMEN_COPY_THREADS");
    const int hw = (int)std::thread::hardware_concurrency();
    const int v = e ? std::atoi(e) : std::min(16, hw > 0 ? hw : 8);
    return v < 1 ? 1 : v > 64 ? 64 : v;
  }();
  const size_t chunk = (bytes + nt - 1) / nt;
  if (nt == 1 || bytes < (8u << 20)) { std::memcpy(dst, src, bytes); return; }
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t) {
    const size_t a = (size_t)t * chunk, b = std::min(bytes, a + chunk);
    if (a >= b) break;
    th.emplace_back([=] { std::memcpy((char*)dst + a, (const char*)src + a, b - a); });
  }
  for (auto& x : th) x.join();
}

// Pageable host -> device through the context's pinned buffers: the host
// copy of this piece overlaps the DMA of the previous one (slot alternates).
void stage_h2d(acu_ctx* c, void* dst_d, const void* src_h, size_t bytes, int& slot) {
  c->pinned(bytes);
  ACU_HIP_CHECK(hipEventSynchronize(c->pin_ev[slot]));  // the buffer's previous DMA is done
  par_copy(c->pin[slot], src_h, bytes);
  ACU_HIP_CHECK(hipMemcpyAsync(dst_d, c->pin[slot], bytes, hipMemcpyHostToDevice, c->stream));
  ACU_HIP_CHECK(hipEventRecord(c->pin_ev[slot], c->stream));
  slot ^= 1;
}

// A pageable host array -> device in pieces of up to 32 MiB through the
// pinned buffers (stage_h2d); the DMAs are queued on the context's stream.
void stage_h2d_pieces(acu_ctx* c, void* dst_d, const void* src_h, size_t bytes) {
  const size_t piece = (size_t)32 << 20;
  int slot = 0;
  for (size_t off = 0; off < bytes; off += piece)
    stage_h2d(c, (char*)dst_d + off, (const char*)src_h + off, std::min(piece, bytes - off), slot);
}

// Device -> pageable host in pieces through the pinned buffers: the DMA of
// one piece overlaps the host copy of the previous one.  Returns when every
// byte is on the host.
void stage_d2h(acu_ctx* c, void* dst_h, const void* src_d, size_t bytes) {
  const size_t piece = std::min(bytes, (size_t)32 << 20);
  if (piece == 0) return;
  c->pinned(piece);
  size_t prev_off = 0, prev_len = 0;
  int prev = -1;
  for (size_t off = 0, k = 0; off < bytes; off += piece, ++k) {
    const size_t len = std::min(piece, bytes - off);
    const int slot = (int)(k & 1);  // its previous piece was copied out in the last iteration
    ACU_HIP_CHECK(hipMemcpyAsync(c->pin[slot], (const char*)src_d + off, len, hipMemcpyDeviceToHost, c->stream));
    ACU_HIP_CHECK(hipEventRecord(c->pin_ev[slot], c->stream));
    if (prev >= 0) {
      ACU_HIP_CHECK(hipEventSynchronize(c->pin_ev[prev]));
      par_copy((char*)dst_h + prev_off, c->pin[prev], prev_len);
    }
    prev = slot; prev_off = off; prev_len = len;
  }
  ACU_HIP_CHECK(hipEventSynchronize(c->pin_ev[prev]));
  par_copy((char*)dst_h + prev_off, c->pin[prev], prev_len);
}

acu_batch* make_volume_batch(acu_ctx* c, int n, int N, int T, int wpu, const float* theta, const acu_batch* like) {
  std::vector<int> Ns(n, N), Ts(n, T), ws(n, wpu);
  std::vector<const float*> th(n, theta);
  acu_batch* b = nullptr;
  check_rc(acu_batch_create(c, n, Ns.data(), Ts.data(), ws.data(), th.data(), &b));
  if (like) {  // same shape: the tuned variants carry over
    b->ramp_variant = like->ramp_variant;
    for (auto& s : b->s) s.fpv = like->s[0].fpv;
    acu::refresh_fp(b);
    if (b->bp_ty != like->bp_ty) acu::set_bp_tile(b, like->bp_ty);
  }
  return b;
}

void volume_run(acu_ctx* ctx, int nslices, int N, int T, int wpu, const float* theta, const float* sino,
                float* fbp, float* tv, float mu, float lambda, int nouter, int ninner, int ncg, int batch,
                int tune, double* stats) {
  const auto t0 = std::chrono::steady_clock::now();
  const int B = std::min(batch, nslices);
  const int groups = (nslices + B - 1) / B;
  const size_t NT = (size_t)N * T, NN = (size_t)N * N;
  Lane lanes[2];
  for (auto& L : lanes) {
    check_rc(acu_ctx_create(ctx->device, &L.ctx));
    L.full = make_volume_batch(L.ctx, B, N, T, wpu, theta, nullptr);
    ACU_HIP_CHECK(hipHostMalloc((void**)&L.sin_pin, (size_t)B * NT * 4, hipHostMallocDefault));
    if (fbp) ACU_HIP_CHECK(hipHostMalloc((void**)&L.fbp_pin, (size_t)B * NN * 4, hipHostMallocDefault));
    if (tv) ACU_HIP_CHECK(hipHostMalloc((void**)&L.tv_pin, (size_t)B * NN * 4, hipHostMallocDefault));
  }
  if (tune) {  // an FBP-only call never runs the projector: its variants stay untimed
    (void)acu::autotune(lanes[0].full, 2, tv ? acu::TUNE_ALL : acu::TUNE_RAMP | acu::TUNE_BP);
    lanes[1].full->ramp_variant = lanes[0].full->ramp_variant;
    for (auto& s : lanes[1].full->s) s.fpv = lanes[0].full->s[0].fpv;
    acu::refresh_fp(lanes[1].full);
    acu::set_bp_tile(lanes[1].full, lanes[0].full->bp_ty);
  }
  const auto t1 = std::chrono::steady_clock::now();
  // results of a lane's last group: wait for its stream, then unstage
  auto finish = [&](Lane& L) {
    if (L.group < 0) return;
    L.ctx->sync();
    const size_t first = (size_t)L.group * B;
    if (fbp) par_copy(fbp + first * NN, L.fbp_pin, (size_t)L.count * NN * 4);
    if (tv) par_copy(tv + first * NN, L.tv_pin, (size_t)L.count * NN * 4);
    L.group = -1;
  };
  for (int g = 0; g < groups; ++g) {
    Lane& L = lanes[g % 2];
    finish(L);
    const int n = std::min(B, nslices - g * B);
    acu_batch* bt = L.full;
    if (n < B) {
      if (!L.tail) L.tail = make_volume_batch(L.ctx, n, N, T, wpu, theta, L.full);
      bt = L.tail;
    }
    par_copy(L.sin_pin, sino + (size_t)g * B * NT, (size_t)n * NT * 4);
    hipStream_t st = L.ctx->stream;
    for (int i = 0; i < n; ++i)
      ACU_HIP_CHECK(hipMemcpyAsync(bt->s[i].f, L.sin_pin + (size_t)i * NT, NT * 4, hipMemcpyHostToDevice, st));
    if (tv) {
      acu::tvopt_run(bt, mu, lambda, nouter, ninner, ncg, ACU_TV_GRAPH);
    } else {
      acu::fbp_all(bt, -1);
      for (auto& s : bt->s) s.fbp_valid = true;
    }
    for (int i = 0; i < n; ++i) {
      if (fbp)
        ACU_HIP_CHECK(hipMemcpyAsync(L.fbp_pin + (size_t)i * NN, bt->s[i].fbp, NN * 4, hipMemcpyDeviceToHost, st));
      if (tv)
        ACU_HIP_CHECK(hipMemcpyAsync(L.tv_pin + (size_t)i * NN, bt->s[i].uk, NN * 4, hipMemcpyDeviceToHost, st));
    }
    L.group = g;
    L.count = n;
  }
  finish(lanes[0]);
  finish(lanes[1]);
  if (stats) {
    const auto t2 = std::chrono::steady_clock::now();
    const double ms = std::chrono::duration<double, std::milli>(t2 - t0).count();
    const double setup = std::chrono::duration<double, std::milli>(t1 - t0).count();
    stats[0] = ms;
    stats[1] = nslices / (ms * 1e-3);
    stats[2] = setup;
    stats[3] = nslices / ((ms - setup) * 1e-3);
  }
}
}  // namespace

int acu_volume_reconstruct(acu_ctx* ctx, int nslices, int nimage, int ntheta, int wpu, const float* theta_deg,
                           const float* sino_host, float* fbp_host, float* tv_host, float mu, float lambda,
                           int nouter, int ninner, int ncg, int batch, int tune, double* stats) {
  if (!ctx || nslices < 1 || !theta_deg || !sino_host || batch < 1 || (!fbp_host && !tv_host))
    return fail(ACU_ERR_ARG, "acu_volume_reconstruct: bad argument");
  if (tv_host && (nouter < 1 || lambda == 0.0f))
    return fail(ACU_ERR_ARG, "acu_volume_reconstruct: TV needs nouter >= 1 and lambda != 0");
  ACU_TRY(hipSetDevice(ctx->device);
          volume_run(ctx, nslices, nimage, ntheta, wpu, theta_deg, sino_host, fbp_host, tv_host, mu, lambda,
                     nouter, ninner, ncg, batch, tune, stats));
}

// ---- projection -> sinogram preprocessing (projections_to_sinogram_hpc.F90) ----
namespace {
// device scratch of the mirror-error sums: block partials + the total
struct PrepScratch {
  double* d = nullptr;
  size_t n = 0;
  ~PrepScratch() { if (d) hipFree(d); }
  double* get(size_t need) {
    if (need > n) {
      if (d) ACU_HIP_CHECK(hipFree(d));
      d = nullptr;
      ACU_HIP_CHECK(hipMalloc(&d, need * sizeof(double)));
      n = need;
    }
    return d;
  }
};

double mirror_error(acu_ctx* c, PrepScratch& sc, int n, int add_nx, float theta, float shift, const float* p0,
                    const float* p180) {
  const acu::PrepGeom g = acu::prep_geom(n, add_nx, theta, shift);
  const int nb = acu::mirror_blocks(g);
  double* d = sc.get((size_t)nb + 1);
  acu::launch_mirror(g, p0, p180, d, d + nb, c->stream);
  ACU_HIP_CHECK(hipGetLastError());
  double e = 0.0;
  ACU_HIP_CHECK(hipMemcpyAsync(&e, d + nb, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  ACU_HIP_CHECK(hipStreamSynchronize(c->stream));
  return e;
}

// First shift estimate from the centroids of the 0 and 180 degree
// projections (:231-248), FP32 weights, FP64 sums.
float centroid_shift(int n, int add_nx, const float* p0, const float* p180) {
#pragma clang fp contract(off)
  const int nw = n + add_nx;
  double tm = 0.0, c0 = 0.0, c180 = 0.0;
  for (int jx = 1; jx <= n; ++jx) {
    const float w = (2.0f * (float)nw * (float)(jx - 1)) / (float)(nw - 1) - (float)n - (float)add_nx;
    for (int jy = 1; jy <= n; ++jy) {
      const size_t i = (size_t)(jx - 1) + (size_t)(jy - 1) * n;
      tm += (double)p0[i] + (double)p180[i];
      c0 += (double)p0[i] * (double)w;
      c180 += (double)p180[i] * (double)w;
    }
  }
  return -((float)(c0 / tm) + (float)(c180 / tm)) / (2.0f * (float)n);
}

// The pattern search of :250-364: FP32 parameters and steps as in the
// reference, each trial's mirror error on the GPU (FP64 sums).
void find_center(acu_ctx* c, int n, int add_nx, const float* p0h, const float* p180h, float* theta_out,
                 float* shift_out, int* evals_out) {
#pragma clang fp contract(off)
  PrepScratch sc;
  float *p0 = nullptr, *p180 = nullptr;
  const size_t bytes = (size_t)n * n * 4;
  struct Free { float*& a; float*& b; ~Free() { if (a) hipFree(a); if (b) hipFree(b); } } fr{p0, p180};
  ACU_HIP_CHECK(hipMalloc(&p0, bytes));
  ACU_HIP_CHECK(hipMalloc(&p180, bytes));
  ACU_HIP_CHECK(hipMemcpyAsync(p0, p0h, bytes, hipMemcpyHostToDevice, c->stream));
  ACU_HIP_CHECK(hipMemcpyAsync(p180, p180h, bytes, hipMemcpyHostToDevice, c->stream));
  float best_shift = centroid_shift(n, add_nx, p0h, p180h), best_theta = 0.0f;
  float dshift = 1e-4f, dtheta = 1e-4f;
  const int n_sp = 6;
  const float n_mag = 1.05f, n_min = 2.0f / (float)n_sp;
  int evals = 1;
  double best = mirror_error(c, sc, n, add_nx, best_theta, best_shift, p0, p180);
  while (std::max(dtheta, dshift) > 1e-8f) {
    double prev = best;
    for (int jsp = 1; jsp <= n_sp; ++jsp) {
      const float th = best_theta + dtheta * (float)(jsp - n_sp / 2);
      const double e = mirror_error(c, sc, n, add_nx, th, best_shift, p0, p180);
      ++evals;
      if (e < best) { best_theta = th; best = e; }
    }
    dtheta = best < prev ? n_mag * dtheta : n_min * dtheta;
    prev = best;
    for (int jsp = 1; jsp <= n_sp; ++jsp) {
      const float sh = best_shift + dshift * (float)(jsp - n_sp / 2);
      const double e = mirror_error(c, sc, n, add_nx, best_theta, sh, p0, p180);
      ++evals;
      if (e < best) { best_shift = sh; best = e; }
    }
    dshift = best < prev ? n_mag * dshift : n_min * dshift;
  }
  *theta_out = best_theta;
  *shift_out = best_shift;
  if (evals_out) *evals_out = evals;
}

// Median filter + correction of every projection into the sinogram volume
// (:384-487): chunks of up to B projections go up with one neighbour on
// each side (the first and last projections are their own neighbours).
void prep_sinograms(acu_ctx* c, int n, int nfiles, const float* proj, float theta, float shift, float* sino_h) {
  const int add_nx = n % 2 == 0 ? 1 : 0, nw = n + add_nx;
  const size_t nn = (size_t)n * n;
  const int B = std::max(1, std::min(nfiles, (int)std::max<size_t>(1, ((size_t)1 << 30) / (nn * 4))));
  float *win = nullptr, *med = nullptr, *sino = nullptr;
  struct Free {
    float*& a; float*& b; float*& s;
    ~Free() { if (a) hipFree(a); if (b) hipFree(b); if (s) hipFree(s); }
  } fr{win, med, sino};
  ACU_HIP_CHECK(hipMalloc(&win, (size_t)(B + 2) * nn * 4));
  ACU_HIP_CHECK(hipMalloc(&med, (size_t)B * nn * 4));
  ACU_HIP_CHECK(hipMalloc(&sino, (size_t)n * nfiles * nw * 4));
  const acu::PrepGeom g = acu::prep_geom(n, add_nx, theta, shift);
  for (int j0 = 0; j0 < nfiles; j0 += B) {
    const int cnt = std::min(B, nfiles - j0);
    // window slot w holds projection j0 - 1 + w (clamped to the range)
    int slot = 0;
    for (int w = 0; w < cnt + 2; ++w) {
      const int j = std::min(std::max(j0 - 1 + w, 0), nfiles - 1);
      stage_h2d(c, win + (size_t)w * nn, proj + (size_t)j * nn, nn * 4, slot);
    }
    acu::timed(c, acu::K_PREP, [&] {
      for (int k = 0; k < cnt; ++k)
        acu::launch_median27(n, win + (size_t)k * nn, win + (size_t)(k + 1) * nn, win + (size_t)(k + 2) * nn,
                             med + (size_t)k * nn, c->stream);
      acu::launch_correct(g, med, cnt, j0, nfiles, sino, c->stream);
    });
    ACU_HIP_CHECK(hipGetLastError());
  }
  stage_d2h(c, sino_h, sino, (size_t)n * nfiles * nw * 4);
  c->sync();
}
}  // namespace

int acu_prep_median27_dev(acu_ctx* ctx, int n, const float* l, const float* c, const float* r, float* out) {
  if (!ctx || n < 1 || !l || !c || !r || !out) return fail(ACU_ERR_ARG, "acu_prep_median27_dev: bad argument");
  ACU_TRY(hipSetDevice(ctx->device); acu::launch_median27(n, l, c, r, out, ctx->stream);
          ACU_HIP_CHECK(hipGetLastError()));
}

int acu_prep_correct_dev(acu_ctx* ctx, int n, int add_nx, float theta, float shift, const float* imgs, int count,
                         int job0, int nfiles, float* sino) {
  if (!ctx || n < 1 || (add_nx != 0 && add_nx != 1) || !imgs || !sino || count < 1 || job0 < 0 ||
      job0 + count > nfiles)
    return fail(ACU_ERR_ARG, "acu_prep_correct_dev: bad argument");
  ACU_TRY(hipSetDevice(ctx->device);
          acu::launch_correct(acu::prep_geom(n, add_nx, theta, shift), imgs, count, job0, nfiles, sino, ctx->stream);
          ACU_HIP_CHECK(hipGetLastError()));
}

int acu_prep_mirror_error_dev(acu_ctx* ctx, int n, int add_nx, float theta, float shift, const float* p0,
                              const float* p180, double* err) {
  if (!ctx || n < 1 || (add_nx != 0 && add_nx != 1) || !p0 || !p180 || !err)
    return fail(ACU_ERR_ARG, "acu_prep_mirror_error_dev: bad argument");
  ACU_TRY(hipSetDevice(ctx->device); PrepScratch sc;
          *err = mirror_error(ctx, sc, n, add_nx, theta, shift, p0, p180));
}

int acu_prep_find_center(acu_ctx* ctx, int n, const float* p0, const float* p180, float* theta, float* shift,
                         int* evaluations) {
  if (!ctx || n < 2 || !p0 || !p180 || !theta || !shift) return fail(ACU_ERR_ARG, "acu_prep_find_center: bad argument");
  ACU_TRY(hipSetDevice(ctx->device); find_center(ctx, n, n % 2 == 0 ? 1 : 0, p0, p180, theta, shift, evaluations));
}

int acu_prep_sinograms(acu_ctx* ctx, int n, int nfiles, const float* projections, float theta, float shift,
                       float* sinograms) {
  if (!ctx || n < 2 || nfiles < 1 || !projections || !sinograms)
    return fail(ACU_ERR_ARG, "acu_prep_sinograms: bad argument");
  ACU_TRY(hipSetDevice(ctx->device); prep_sinograms(ctx, n, nfiles, projections, theta, shift, sinograms));
}

int acu_ingest_mean_dev(acu_ctx* ctx, int count, long long npix, const float* imgs, float* out) {
  if (!ctx || count < 1 || npix < 1 || !imgs || !out) return fail(ACU_ERR_ARG, "acu_ingest_mean_dev: bad argument");
  ACU_TRY(hipSetDevice(ctx->device); acu::timed(ctx, acu::K_INGEST, [&] { acu::launch_mean(imgs, count, (size_t)npix, out, ctx->stream); });
          ACU_HIP_CHECK(hipGetLastError()));
}

int acu_ingest_normalize_dev(acu_ctx* ctx, int count, long long npix, const float* imgs, const float* open_beam,
                             float* out) {
  if (!ctx || count < 1 || npix < 1 || !imgs || !open_beam || !out)
    return fail(ACU_ERR_ARG, "acu_ingest_normalize_dev: bad argument");
  ACU_TRY(hipSetDevice(ctx->device); acu::timed(ctx, acu::K_INGEST,
                     [&] { acu::launch_normalize(imgs, count, (size_t)npix, open_beam, out, ctx->stream); });
          ACU_HIP_CHECK(hipGetLastError()));
}

int acu_ingest_mean(acu_ctx* ctx, int count, long long npix, const float* imgs, float* out) {
  if (!ctx || count < 1 || npix < 1 || !imgs || !out) return fail(ACU_ERR_ARG, "acu_ingest_mean: bad argument");
  ACU_TRY(hipSetDevice(ctx->device);
          const size_t np = (size_t)npix;
          DevBuf in(np * count), o(np);
          stage_h2d_pieces(ctx, in.p, imgs, np * count * 4);
          acu::timed(ctx, acu::K_INGEST, [&] { acu::launch_mean(in.p, count, np, o.p, ctx->stream); }); ACU_HIP_CHECK(hipGetLastError());
          stage_d2h(ctx, out, o.p, np * 4);
          ctx->sync());
}

int acu_ingest_normalize(acu_ctx* ctx, int count, long long npix, const float* imgs, const float* open_beam,
                         float* out) {
  if (!ctx || count < 1 || npix < 1 || !imgs || !open_beam || !out)
    return fail(ACU_ERR_ARG, "acu_ingest_normalize: bad argument");
  ACU_TRY(hipSetDevice(ctx->device);
          const size_t np = (size_t)npix;
          DevBuf in(np * count), ob(np), o(np * count);
          stage_h2d_pieces(ctx, in.p, imgs, np * count * 4);
          stage_h2d_pieces(ctx, ob.p, open_beam, np * 4);
          acu::timed(ctx, acu::K_INGEST, [&] { acu::launch_normalize(in.p, count, np, ob.p, o.p, ctx->stream); }); ACU_HIP_CHECK(hipGetLastError());
          stage_d2h(ctx, out, o.p, np * count * 4);
          ctx->sync());
}

int acu_tvopt(acu_batch* b, const float* f_host, float mu, float lambda, int nouter, int ninner,
              int ncg, float* img_tv_host, float* img_fbp_host, float* delta_host) {
  int rc = acu_batch_set_sino(b, 0, f_host);
  if (rc) return rc;
  if ((rc = acu_batch_tvopt(b, mu, lambda, nouter, ninner, ncg, ACU_TV_GRAPH))) return rc;
  if (img_tv_host && (rc = acu_batch_get_image(b, 0, ACU_IMG_TV, img_tv_host))) return rc;
  if (img_fbp_host && (rc = acu_batch_get_image(b, 0, ACU_IMG_FBP, img_fbp_host))) return rc;
  if (delta_host && nouter > 0 && (rc = acu_batch_get_delta(b, 0, delta_host, nouter))) return rc;
  return ACU_OK;
}

}  // extern "C"
