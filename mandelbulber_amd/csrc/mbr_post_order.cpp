// mbr_post_order.cpp — host side of the depth-of-field post effect: the
// order in which PostRendering_DOF (src/image.cpp:57-247) splats its pixels.
//
// The reference sorts the zBuffer (QuickSortZBuffer, src/common_math.h:53-84,
// called on indices [1, n-1]) and then, from the far end, swaps each entry
// with a randomly drawn earlier one of similar blur size (image.cpp:96-152).
// Both steps are sequential and define the result (splats blend in this
// order), so they run here, on the host, exactly; the splatting itself runs
// on the GPU as an order-preserving gather (mbr_post.hip).
#include <cstdint>
#include <vector>

#include "mbr_kernels.hpp"
#include "mbr_rand.hpp"

namespace mbr {

namespace {

struct ZEntry {
  float z;
  int i;
};

// Hoare partition around the middle element, ascending z; the two sides are
// sorted from an explicit stack (the order of independent ranges does not
// change the result).
void zsort(std::vector<ZEntry> &d, int lo, int hi) {
  std::vector<std::pair<int, int>> stack;
  stack.push_back({lo, hi});
  while (!stack.empty()) {
    const int l = stack.back().first, p = stack.back().second;
    stack.pop_back();
    int i = l, j = p;
    const ZEntry x = d[(l + p) / 2];
    do {
      while (d[i].z < x.z) i++;
      while (x.z < d[j].z) j--;
      if (i <= j) {
        const ZEntry t = d[i];
        d[i] = d[j];
        d[j] = t;
        i++;
        j--;
      }
    } while (i <= j);
    if (l < j) stack.push_back({l, j});
    if (i < p) stack.push_back({i, p});
  }
}

}  // namespace

void dof_order(const float *z, int width, int height, double neutral, double deep, uint32_t seed, uint32_t *rank) {
  const int n = width * height;
  std::vector<ZEntry> d((size_t)n);
  for (int k = 0; k < n; k++) d[k] = {z[k], k};
  if (n > 2) zsort(d, 1, n - 1);
  GlibcRand rnd(seed);
  for (int i = n - 1; i >= 0; i--) {
    const ZEntry t = d[i];
    const double z1 = t.z;
    const double size1 = (z1 - neutral) / z1 * deep;
    int step = i, ii;
    bool done;
    do {
      ii = i - rnd.next() % (step + 1);
      if (ii <= 0) ii = 0;
      const double z2 = d[ii].z;
      const double size2 = (z2 - neutral) / z2 * deep;
      // (the reference's "both sizes 0: done" is overwritten by this test)
      if (size1 * size2 > 0) {
        const double cmp = size1 > 0 ? size2 / size1 : size1 / size2;
        done = cmp > 0.7;
      } else {
        done = false;
      }
      step = (int)(step * 0.7 - 1);
      if (step <= 0) done = true;
    } while (!done);
    d[i] = d[ii];
    d[ii] = t;
  }
  // the splat loop runs from the end of the array
  for (int k = 0; k < n; k++) rank[d[(size_t)n - 1 - k].i] = (uint32_t)k;
}

}  // namespace mbr
