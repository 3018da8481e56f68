// median_net.hpp -- a comparator network that selects the 14th smallest of
// 27 values (the median of a 3x3x3 neighbourhood), built at compile time by
// forgetful selection:
//   keep a working set of (27 + 3) / 2 = 15 values; while values remain
//   unseen, move the set's minimum and maximum out of it with
//   compare-exchanges (pair the set up, then a minimum tournament over the
//   pairs' smaller and a maximum tournament over their larger elements) --
//   with 15 values in the set the minimum has 14 larger and the maximum 14
//   smaller values, so neither is the median, and the median of the rest
//   is the median of the whole -- and bring the next value in; the last
//   three values' median is the answer.
// 153 compare-exchanges (306 min/max), against 240 for a 32-value bitonic
// sort.  Plain C++ (no HIP), so that a host program can print the network
// for an exhaustive check by the 0-1 principle.
#pragma once

namespace acu {

struct MedianNet {
  int n = 0;                   // compare-exchanges
  unsigned char lo[160] = {};  // after exchange i: slot lo[i] holds the smaller value,
  unsigned char hi[160] = {};  // slot hi[i] the larger
  int out = 0;                 // slot of the median at the end
};

constexpr MedianNet make_median27_net() {
  MedianNet net;
  int set[27] = {};
  int size = 15, next = 15;
  for (int i = 0; i < size; ++i) set[i] = i;
  auto ce = [&net](int a, int b) {
    net.lo[net.n] = (unsigned char)a;
    net.hi[net.n] = (unsigned char)b;
    ++net.n;
  };
  while (!(size == 3 && next == 27)) {
    int lo[8] = {}, hi[8] = {}, nl = 0, nh = 0;
    for (int i = 0; i + 1 < size; i += 2) {
      ce(set[i], set[i + 1]);
      lo[nl++] = set[i];
      hi[nh++] = set[i + 1];
    }
    if (size % 2) {
      lo[nl++] = set[size - 1];
      hi[nh++] = set[size - 1];
    }
    const int mn = lo[0];
    for (int i = 1; i < nl; ++i) ce(mn, lo[i]);  // mn keeps the smallest
    const int mx = hi[0];
    for (int i = 1; i < nh; ++i)
      if (hi[i] != mn) ce(hi[i], mx);           // mx keeps the largest
    int k = 0;
    for (int i = 0; i < size; ++i)
      if (set[i] != mn && set[i] != mx) set[k++] = set[i];
    set[k++] = next++;
    size = k;
  }
  ce(set[0], set[1]);
  ce(set[1], set[2]);
  ce(set[0], set[1]);
  net.out = set[1];
  return net;
}

constexpr MedianNet MEDIAN27_NET = make_median27_net();
static_assert(MEDIAN27_NET.n == 153, "forgetful selection of the median of 27");

}  // namespace acu
