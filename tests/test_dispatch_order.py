"""The LDS projector's dispatch order (fpl_dispatch_order in
acumen_amd/csrc/fpl.hip) and the backprojector's XCD shares (bp_xcd_shares in
bp.hip), checked on the CPU through the built library.  The order, for single
slices, the 4-slice 2049^2 x 1500 launch and random mixed batches, is empty
(the natural order, kept when its XCD shares are balanced, as with 8 equal
slices) or a permutation of the launch's blocks (a wrong index would drop or repeat ray
blocks), and the angle blocks of a ray block stay consecutive and in order
within an XCD's share.  The shares cover the launch's tiles in order, each
within one tile's work of an eighth of the total, and 8 equal slices give
one slice per XCD.  The pitch classes of the projector's angle blocks
(fpl_pitch_classes) are checked for range, extent and determinism."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "acumen_amd")

CHECKER = r"""
#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>
namespace acu {
struct FplGeom { int N, nab, block_begin; };
std::vector<int> fpl_dispatch_order(const std::vector<FplGeom>& g, int base, int n);
struct BPShares { int b[9]; int C; };
BPShares bp_xcd_shares(const std::vector<float>& w);
}
static int check_shares(const std::vector<float>& w) {
  const acu::BPShares sh = acu::bp_xcd_shares(w);
  const int n = (int)w.size();
  if (sh.b[0] != 0 || sh.b[8] != n) return 11;
  double tot = 0, mx = 0;
  for (float v : w) { tot += v; mx = std::max(mx, (double)v); }
  int C = 1;
  for (int k = 0; k < 8; ++k) {
    if (sh.b[k + 1] < sh.b[k]) return 12;
    C = std::max(C, sh.b[k + 1] - sh.b[k]);
    double s = 0;
    for (int i = sh.b[k]; i < sh.b[k + 1]; ++i) s += w[i];
    if (s > tot / 8 + mx + 1e-6 * tot) return 13;
  }
  return C == sh.C ? 0 : 14;
}
using acu::FplGeom;
static int blocks(const FplGeom& t) { return (t.N + 63) / 64 * t.nab; }
// the launch [base, base + n) covers whole slices g
static int check(const std::vector<FplGeom>& g, int base, int n) {
  std::vector<int> o = acu::fpl_dispatch_order(g, base, n);
  if (o.empty())  // the natural order
    for (int i = 0; i < n; ++i) o.push_back(i);
  if ((int)o.size() != n) return 1;
  std::vector<int> seen(n, 0);
  for (int v : o) {
    if (v < 0 || v >= n) return 2;
    if (seen[v]++) return 3;
  }
  // segment (slice, ray block) of each block; within a share a segment's
  // blocks form one run of consecutive indices in increasing order
  const int C = (n + 7) / 8;
  for (int k = 0; k * C < n; ++k) {
    const int lo = k * C, hi = std::min(n, lo + C);
    std::vector<int> done_seg;
    for (int i = lo; i < hi; ++i) {
      const int gb = base + o[i];
      int t = 0;
      while (t + 1 < (int)g.size() && g[t + 1].block_begin <= gb) ++t;
      const int seg = t * 4096 + (gb - g[t].block_begin) / g[t].nab;
      if (done_seg.empty() || done_seg.back() != seg) {
        for (int s : done_seg) if (s == seg) return 4;  // a ray block split within the share
        done_seg.push_back(seg);
      } else if (o[i] != o[i - 1] + 1) {
        return 5;  // angle blocks out of order
      }
    }
  }
  return 0;
}
int main() {
  std::vector<FplGeom> g;
  int b = 0;
  for (int i = 0; i < 4; ++i) { g.push_back({2049, 94, b}); b += blocks(g.back()); }
  if (int r = check(g, 0, b)) { printf("headline %d\n", r); return 1; }
  // 8 equal slices: one slice per XCD in the natural order, kept; one slice: reordered
  {
    std::vector<FplGeom> g8;
    int b8 = 0;
    for (int i = 0; i < 8; ++i) { g8.push_back({2049, 94, b8}); b8 += blocks(g8.back()); }
    if (!acu::fpl_dispatch_order(g8, 0, b8).empty()) { printf("8 slices reordered\n"); return 1; }
    if (acu::fpl_dispatch_order({g[0]}, 0, blocks(g[0])).empty()) { printf("1 slice kept\n"); return 1; }
  }
  for (int i = 0; i < 4; ++i)
    if (int r = check({g[i]}, g[i].block_begin, blocks(g[i]))) { printf("slice %d: %d\n", i, r); return 1; }
  {
    std::vector<float> w;
    for (int i = 0; i < 8 * 2145; ++i) w.push_back((i % 2145) < 300 ? 30.0f : 1500.0f);
    const acu::BPShares sh = acu::bp_xcd_shares(w);
    for (int k = 0; k <= 8; ++k) if (sh.b[k] != 2145 * k) { printf("shares of 8 slices\n"); return 1; }
    if (int r = check_shares(w)) { printf("shares %d\n", r); return 1; }
  }
  std::mt19937 rng(11);
  for (int n = 1; n <= 20; ++n)
    if (int r = check_shares(std::vector<float>(n, 7.0f))) { printf("small shares %d: %d\n", n, r); return 1; }
  for (int it = 0; it < 500; ++it) {
    std::vector<float> w(1 + rng() % 20000);
    for (auto& v : w) v = (rng() % 4 == 0) ? 1.0f : (float)(360 + rng() % 1141);
    if (int r = check_shares(w)) { printf("random shares %d: %d\n", it, r); return 1; }
  }
  for (int it = 0; it < 2000; ++it) {
    std::vector<FplGeom> m;
    int bb = 0;
    const int ns = 1 + rng() % 10;
    for (int i = 0; i < ns; ++i) { m.push_back({3 + (int)(rng() % 2100), 1 + (int)(rng() % 380), bb}); bb += blocks(m.back()); }
    if (int r = check(m, 0, bb)) { printf("mixed %d: %d\n", it, r); return 1; }
    const int k = rng() % ns;
    if (int r = check({m[k]}, m[k].block_begin, blocks(m[k]))) { printf("mixed slice %d: %d\n", it, r); return 1; }
  }
  printf("ok\n");
  return 0;
}
"""


def test_dispatch_order_is_a_permutation_with_whole_ray_blocks(tmp_path):
    so = os.path.join(LIB, "libacumen_hip.so")
    if not os.path.exists(so):
        pytest.fail("libacumen_hip.so not built (python -c 'import __graft_entry__ as g; g.build()')")
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    src = tmp_path / "order_check.cpp"
    src.write_text(CHECKER)
    exe = str(tmp_path / "order_check")
    subprocess.run([cxx, "-O1", "-std=c++17", str(src), "-L", LIB, "-lacumen_hip", f"-Wl,-rpath,{LIB}", "-o", exe],
                   check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr


PITCH_CHECKER = r"""
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>
namespace acu {
struct Error { std::string msg; };
void fpl_pitch_classes(int N, int T, const float* theta_deg, int al, uint8_t* out);
}
int main() {
  const int Ns[] = {3, 65, 513, 2049, 4097}, Ts[] = {1, 7, 180, 1500};
  for (int N : Ns)
    for (int T : Ts)
      for (int al : {1, 2, 4}) {
        std::vector<float> th(T);
        for (int a = 0; a < T; ++a) th[a] = (float)(a * (180.0 / T));
        const int nb = (T + 4 * al - 1) / (4 * al);
        std::vector<uint8_t> a(nb + 1, 0xEE), b(nb + 1, 0xEE);
        acu::fpl_pitch_classes(N, T, th.data(), al, a.data());
        acu::fpl_pitch_classes(N, T, th.data(), al, b.data());
        if (a[nb] != 0xEE) { printf("wrote past the blocks N=%d T=%d al=%d\n", N, T, al); return 1; }
        for (int k = 0; k < nb; ++k)
          if (a[k] > 7 || a[k] != b[k]) { printf("class N=%d T=%d al=%d k=%d: %d\n", N, T, al, k, a[k]); return 1; }
      }
  std::vector<float> th(8, 1.0f);
  uint8_t out[8];
  try {
    acu::fpl_pitch_classes(513, 8, th.data(), 3, out);
  } catch (const acu::Error&) {
    printf("ok\n");
    return 0;
  }
  printf("al = 3 accepted\n");
  return 1;
}
"""


def test_pitch_classes_in_range_deterministic_and_checked(tmp_path):
    """fpl_pitch_classes (acumen_amd/csrc/fpl.hip), the host bank model that
    picks a row pitch per angle block under the projector's lane map: one
    class in 0..7 per block of 4*al angles, nothing written past the last
    block, the same answer on a second call, and al outside {1, 2, 4}
    rejected."""
    so = os.path.join(LIB, "libacumen_hip.so")
    if not os.path.exists(so):
        pytest.fail("libacumen_hip.so not built (python -c 'import __graft_entry__ as g; g.build()')")
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    src = tmp_path / "pitch_check.cpp"
    src.write_text(PITCH_CHECKER)
    exe = str(tmp_path / "pitch_check")
    subprocess.run([cxx, "-O1", "-std=c++17", str(src), "-L", LIB, "-lacumen_hip", f"-Wl,-rpath,{LIB}", "-o", exe],
                   check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
