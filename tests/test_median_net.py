"""The median-of-27 comparator network of acumen_amd/csrc/median_net.hpp
(used by the preprocessing median filter, projections_to_sinogram_hpc.F90:
420-447): printed by a host program built from the header, then checked on
all 2^27 binary inputs.  By the 0-1 principle a comparator network that
leaves the k-th smallest in a slot for every 0-1 input does so for every
input from a totally ordered set, so this proves it selects the median."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "acumen_amd", "csrc")

PRINTER = r"""
#include <cstdio>
#include "median_net.hpp"
int main() {
  std::printf("%d %d\n", acu::MEDIAN27_NET.n, acu::MEDIAN27_NET.out);
  for (int i = 0; i < acu::MEDIAN27_NET.n; ++i)
    std::printf("%d %d\n", acu::MEDIAN27_NET.lo[i], acu::MEDIAN27_NET.hi[i]);
}
"""


@pytest.fixture(scope="module")
def network(tmp_path_factory):
    d = tmp_path_factory.mktemp("net")
    src, exe = d / "print_net.cpp", d / "print_net"
    src.write_text(PRINTER)
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", CSRC, str(src), "-o", str(exe)], check=True)
    lines = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    n, out = map(int, lines[0].split())
    pairs = [tuple(map(int, l.split())) for l in lines[1:1 + n]]
    return pairs, out


def run_network(pairs, values):
    v = list(values)
    for a, b in pairs:
        v[a], v[b] = min(v[a], v[b]), max(v[a], v[b])
    return v


def test_network_shape(network):
    pairs, out = network
    assert len(pairs) == 153
    assert all(0 <= a < 27 and 0 <= b < 27 and a != b for a, b in pairs)
    assert 0 <= out < 27


def test_selects_median_of_every_binary_input(network):
    """Bit-sliced: input k in [0, 2^27) puts bit j of k in slot j; 64 inputs
    per machine word, min = AND and max = OR; the median of a 0-1 input is 1
    iff it has at least 14 ones."""
    pairs, out = network
    W = 1 << 21
    words = np.arange(W, dtype=np.uint64)
    lane_bits = np.arange(64, dtype=np.uint64)
    slots = []
    for j in range(27):
        if j < 6:
            pat = int(np.bitwise_or.reduce(((lane_bits >> np.uint64(j)) & np.uint64(1)) << lane_bits))
            slots.append(np.full(W, pat, dtype=np.uint64))
        else:
            bit = (words >> np.uint64(j - 6)) & np.uint64(1)
            slots.append(np.where(bit == 1, np.uint64(0xFFFFFFFFFFFFFFFF), np.uint64(0)))
    for a, b in pairs:
        x, y = slots[a], slots[b]
        slots[a], slots[b] = x & y, x | y
    ones_in_lane = np.array([bin(l).count("1") for l in range(64)])
    lanes_at_least = {t: int(sum(1 << l for l in range(64) if ones_in_lane[l] >= t)) for t in range(-7, 22)}
    ones_in_word = np.zeros(W, dtype=np.int64)
    w = words.astype(np.int64)
    while w.any():
        ones_in_word += w & 1
        w >>= 1
    expected = np.array([lanes_at_least[14 - p] for p in range(22)], dtype=np.uint64)[ones_in_word]
    assert np.array_equal(slots[out], expected)


def test_matches_a_sort_on_random_and_tied_values(network):
    pairs, out = network
    rng = np.random.default_rng(5)
    for trial in range(2000):
        vals = rng.integers(0, 4 if trial % 2 else 1000, 27).astype(np.float64)
        assert run_network(pairs, vals)[out] == np.sort(vals)[13]
