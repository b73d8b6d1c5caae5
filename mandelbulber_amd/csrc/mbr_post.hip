// mbr_post.hip — post effects on a rendered frame: screen-space ambient
// occlusion (PostRendering_SSAO / ThreadSSAO, src/image.cpp:249-415) and
// depth of field (PostRendering_DOF, src/image.cpp:57-247).
//
// SSAO is a gather over the zBuffer: every pixel reads up to q(q-1) depths
// at r^2 s (r = 1..q-1, s = width / (2 q^2)) along q screen directions,
// keeps the steepest horizon -dz/dr per direction and adds colour x AO x
// intensity x (1 - opacity) to its own compiled 16-bit value.  No pixel
// reads another pixel's output, so one lane per pixel needs no ordering.
// The arithmetic is the reference's FP64 in the reference's order with
// contraction off; the direction table comes from the host's libm.
// Workgroups cover 16x16 pixel tiles so the near samples of a tile share
// cache lines; the far samples (up to width/2 away) come from L2.
#include <hipcub/hipcub.hpp>

#include "mbr_kernels.hpp"

namespace mbr {

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kTile = 16;

// Plane position of the centre pixel (src/image.cpp:288-314).
__device__ __forceinline__ void ssao_centre(const SsaoParams &p, int x, int y, double z, double &px, double &py) {
#pragma clang fp contract(off)
  double x2, y2;
  if (p.type == 1) {  // fish eye
    x2 = kPi * ((double)x / p.width - 0.5) * p.aspect;
    y2 = kPi * ((double)y / p.height - 0.5);
    const double r = sqrt(x2 * x2 + y2 * y2);
    if (r != 0.0) {
      x2 = x2 / r * sin(r * p.fov) * z;
      y2 = y2 / r * sin(r * p.fov) * z;
    }
  } else if (p.type == 2) {  // equirectangular
    x2 = kPi * ((double)x / p.width - 0.5) * p.aspect;
    y2 = kPi * ((double)y / p.height - 0.5);
    x2 = sin(p.fov * x2) * cos(p.fov * y2) * z;
    y2 = sin(p.fov * y2) * z;
  } else {
    x2 = ((double)x / p.width - 0.5) * p.aspect;
    y2 = ((double)y / p.height - 0.5);
    x2 = x2 * z * p.fov;
    y2 = y2 * z * p.fov;
  }
  px = x2;
  py = y2;
}

// Plane position of a sample (src/image.cpp:337-362): its own grouping of
// z2 * persp, and in the fish eye the guard tests the loop radius (never 0).
__device__ __forceinline__ void ssao_sample(const SsaoParams &p, double xx, double yy, double z2, double &px,
                                            double &py) {
#pragma clang fp contract(off)
  double xx2, yy2;
  if (p.type == 1) {
    xx2 = kPi * (xx / p.width - 0.5) * p.aspect;
    yy2 = kPi * (yy / p.height - 0.5);
    const double r2 = sqrt(xx2 * xx2 + yy2 * yy2);
    xx2 = xx2 / r2 * sin(r2 * p.fov) * z2;
    yy2 = yy2 / r2 * sin(r2 * p.fov) * z2;
  } else if (p.type == 2) {
    xx2 = kPi * (xx / p.width - 0.5) * p.aspect;
    yy2 = kPi * (yy / p.height - 0.5);
    xx2 = sin(p.fov * xx2) * cos(p.fov * yy2) * z2;
    yy2 = sin(p.fov * yy2) * z2;
  } else {
    xx2 = (xx / p.width - 0.5) * p.aspect;
    yy2 = (yy / p.height - 0.5);
    xx2 = xx2 * (z2 * p.fov);
    yy2 = yy2 * (z2 * p.fov);
  }
  px = xx2;
  py = yy2;
}

__global__ void __launch_bounds__(kTile *kTile)
    k_ssao(SsaoParams p, const float *__restrict__ zb, const uint16_t *__restrict__ opacity,
           const uint8_t *__restrict__ cr, const uint8_t *__restrict__ cg, const uint8_t *__restrict__ cb,
           uint16_t *__restrict__ image16) {
#pragma clang fp contract(off)
  const int x = blockIdx.x * kTile + (threadIdx.x % kTile);
  const int y = blockIdx.y * kTile + (threadIdx.x / kTile);
  if (x >= p.width || y >= p.height) return;
  const size_t i = (size_t)y * p.width + x;
  const double z = zb[i];
  const double op = opacity[i] / 65535.0;
  double total = 0;
  if (z < 1e19) {
    double x2, y2;
    ssao_centre(p, x, y, z, x2, y2);
    double ambient = 0;
    for (int a = 0; a < p.quality; a++) {
      const double ca = p.cs[a], sa = p.sn[a];
      double best = -1e50;  // steepest -dz/dr along this direction
      for (double r = 1.0; r < p.quality; r += 1.0) {
        const double rr = r * r * p.scale;
        const double xx = x + rr * ca, yy = y + rr * sa;
        if ((int)xx == x && (int)yy == y) continue;
        if (xx < 0 || xx > p.width - 1 || yy < 0 || yy > p.height - 1) continue;
        const double z2 = zb[(size_t)(int)yy * p.width + (int)xx];
        double xx2, yy2;
        ssao_sample(p, xx, yy, z2, xx2, yy2);
        const double dx = xx2 - x2, dy = yy2 - y2, dz = z2 - z;
        const double dr = sqrt(dx * dx + dy * dy);
        const double diff = -dz / dr;
        if (diff > best) best = diff;
      }
      ambient += -atan(best) / kPi + 0.5;
    }
    total = ambient / p.quality;
    if (total < 0) total = 0;
  }
  const uint8_t col[3] = {cr[i], cg[i], cb[i]};
  for (int c = 0; c < 3; c++) {  // src/image.cpp:375-383
    double v = image16[3 * i + c] + 65535.0 * col[c] / 256.0 * total * p.intensity * (1.0 - op);
    if (v > 65535) v = 65535;
    image16[3 * i + c] = (uint16_t)v;
  }
}

// ------------------------------------------------------------------ DOF
//
// The reference splats every pixel, in an order fixed on the host (z sort,
// then a randomised shuffle among pixels of similar blur), as a disc of
// radius blur = |z - focus| / z * deep (at most 100) with the weight
// op = (blur - r) / (blur^2.5 pi / 3), clamped to [0, 1]; each splat blends
// the pixel's ORIGINAL colour into the CURRENT image, truncating to 16 bits
// (image.cpp:165-205).  The result at one pixel is therefore the fold of
// the splats that cover it, in splat order, starting from its own original
// value.  This gather reproduces it exactly and without write races:
//   1. every source lists the 16x16 tiles its disc's bounding box touches
//      (key = tile, then splat rank);
//   2. a radix sort puts each tile's sources in splat order;
//   3. one workgroup per tile streams its list through LDS, and every
//      pixel folds the entries that cover it, in order.
// Work is the sum over tiles of their list lengths times 256 cheap tests;
// the blends themselves equal the reference's.

__device__ __forceinline__ double dof_blur(const DofParams &p, float zf) {
#pragma clang fp contract(off)
  const double z = zf;
  double blur = fabs(z - p.neutral) / z * p.deep;
  if (blur > 100) blur = 100.0;
  return blur;
}

// Tile bounding box of source s's splat (the reference's loop bounds
// x - size .. x + size clipped to the image); false when the splat cannot
// change any pixel (blur not > 0: the weight is 0 or undefined).
__device__ __forceinline__ bool dof_box(const DofParams &p, int s, float zf, int &tx0, int &tx1, int &ty0,
                                        int &ty1) {
  const double blur = dof_blur(p, zf);
  if (!(blur > 0.0)) return false;
  const int size = (int)blur;
  const int x = s % p.width, y = s / p.width;
  tx0 = max(0, x - size) / kTile;
  tx1 = min(p.width - 1, x + size) / kTile;
  ty0 = max(0, y - size) / kTile;
  ty1 = min(p.height - 1, y + size) / kTile;
  return true;
}

__global__ void __launch_bounds__(256) k_dof_count(DofParams p, const float *__restrict__ zb,
                                                   unsigned long long *__restrict__ cnt) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= p.width * p.height) return;
  int tx0, tx1, ty0, ty1;
  cnt[s] = dof_box(p, s, zb[s], tx0, tx1, ty0, ty1) ? (unsigned long long)(tx1 - tx0 + 1) * (ty1 - ty0 + 1) : 0ull;
}

__global__ void __launch_bounds__(256) k_dof_fill(DofParams p, const float *__restrict__ zb,
                                                  const uint32_t *__restrict__ rank,
                                                  const unsigned long long *__restrict__ off,
                                                  unsigned long long *__restrict__ keys, uint32_t *__restrict__ vals) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= p.width * p.height) return;
  int tx0, tx1, ty0, ty1;
  if (!dof_box(p, s, zb[s], tx0, tx1, ty0, ty1)) return;
  unsigned long long o = off[s];
  const unsigned long long r = rank[s];
  for (int ty = ty0; ty <= ty1; ty++)
    for (int tx = tx0; tx <= tx1; tx++, o++) {
      keys[o] = ((unsigned long long)(ty * p.tiles_x + tx) << p.rank_bits) | r;
      vals[o] = (uint32_t)s;
    }
}

// seg[t] = first list entry of tile t (lower bound of its key range);
// seg[tiles] = the total.
__global__ void __launch_bounds__(256) k_dof_segments(DofParams p, const unsigned long long *__restrict__ keys,
                                                      long long n, unsigned long long *__restrict__ seg) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  const int tiles = p.tiles_x * p.tiles_y;
  if (t > tiles) return;
  if (t == tiles) {
    seg[t] = (unsigned long long)n;
    return;
  }
  const unsigned long long key = (unsigned long long)t << p.rank_bits;
  long long lo = 0, hi = n;
  while (lo < hi) {
    const long long mid = (lo + hi) / 2;
    if (keys[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  seg[t] = (unsigned long long)lo;
}

__global__ void __launch_bounds__(kTile *kTile)
    k_dof_fold(DofParams p, const float *__restrict__ zb, const unsigned long long *__restrict__ seg,
               const uint32_t *__restrict__ vals, const uint16_t *__restrict__ img_in,
               const uint16_t *__restrict__ alpha_in, uint16_t *__restrict__ img_out,
               uint16_t *__restrict__ alpha_out) {
#pragma clang fp contract(off)
  __shared__ int s_x[kTile * kTile], s_y[kTile * kTile], s_size[kTile * kTile];
  __shared__ double s_blur[kTile * kTile], s_factor[kTile * kTile];
  __shared__ uint16_t s_c[4][kTile * kTile];
  const int t = blockIdx.x;
  const int tx = t % p.tiles_x, ty = t / p.tiles_x;
  const int px = tx * kTile + (threadIdx.x % kTile), py = ty * kTile + (threadIdx.x / kTile);
  const bool inside = px < p.width && py < p.height;
  const size_t q = inside ? (size_t)py * p.width + px : 0;
  uint16_t v[4] = {0, 0, 0, 0};
  if (inside) {
    v[0] = img_in[3 * q];
    v[1] = img_in[3 * q + 1];
    v[2] = img_in[3 * q + 2];
    v[3] = alpha_in ? alpha_in[q] : 0;
  }
  const unsigned long long end = seg[t + 1];
  for (unsigned long long start = seg[t]; start < end; start += kTile * kTile) {
    const unsigned long long e = start + threadIdx.x;
    if (e < end) {
      const int s = (int)vals[e];
      const double blur = dof_blur(p, zb[s]);
      s_x[threadIdx.x] = s % p.width;
      s_y[threadIdx.x] = s / p.width;
      s_size[threadIdx.x] = (int)blur;
      s_blur[threadIdx.x] = blur;
      s_factor[threadIdx.x] = blur * blur * sqrt(blur) * kPi / 3.0;
      s_c[0][threadIdx.x] = img_in[3 * (size_t)s];
      s_c[1][threadIdx.x] = img_in[3 * (size_t)s + 1];
      s_c[2][threadIdx.x] = img_in[3 * (size_t)s + 2];
      s_c[3][threadIdx.x] = alpha_in ? alpha_in[s] : 0;
    }
    __syncthreads();
    const int cnt = (int)min((unsigned long long)(kTile * kTile), end - start);
    if (inside) {
      for (int k = 0; k < cnt; k++) {
        const int dx = px - s_x[k], dy = py - s_y[k], size = s_size[k];
        if (dx < -size || dx > size || dy < -size || dy > size) continue;
        // sqrt of the float (dx^2 + dy^2) correctly rounded to float (the
        // reference's float overload): the FP64 root rounded to FP32 is
        // that value (53 >= 2 x 24 + 2 bits)
        const double r = (double)(float)sqrt((double)(dx * dx + dy * dy));
        double op = (s_blur[k] - r) / s_factor[k];
        if (op > 1.0) op = 1.0;
        if (op < 0.0) op = 0.0;
        if (op > 0.0) {
          const double opN = 1.0 - op;
          for (int c = 0; c < 4; c++) v[c] = (uint16_t)(v[c] * opN + s_c[c][k] * op);
        }
      }
    }
    __syncthreads();
  }
  if (inside) {
    img_out[3 * q] = v[0];
    img_out[3 * q + 1] = v[1];
    img_out[3 * q + 2] = v[2];
    if (alpha_out) alpha_out[q] = v[3];
  }
}

}  // namespace

hipError_t dof_count(const DofParams &p, const float *z, unsigned long long *cnt, hipStream_t s) {
  const int n = p.width * p.height;
  hipLaunchKernelGGL(k_dof_count, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, z, cnt);
  return hipGetLastError();
}

hipError_t dof_scan(void *tmp, size_t &tmp_bytes, const unsigned long long *cnt, unsigned long long *off, int n,
                    hipStream_t s) {
  return hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, cnt, off, n, s);
}

hipError_t dof_fill(const DofParams &p, const float *z, const uint32_t *rank, const unsigned long long *off,
                    unsigned long long *keys, uint32_t *vals, hipStream_t s) {
  const int n = p.width * p.height;
  hipLaunchKernelGGL(k_dof_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, z, rank, off, keys, vals);
  return hipGetLastError();
}

hipError_t dof_sort(void *tmp, size_t &tmp_bytes, const unsigned long long *kin, unsigned long long *kout,
                    const uint32_t *vin, uint32_t *vout, int n, int bits, hipStream_t s) {
  return hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, kin, kout, vin, vout, n, 0, bits, s);
}

hipError_t dof_fold(const DofParams &p, const float *z, const unsigned long long *keys, long long n,
                    unsigned long long *seg, const uint32_t *vals, const uint16_t *img_in, const uint16_t *alpha_in,
                    uint16_t *img_out, uint16_t *alpha_out, hipStream_t s) {
  const int tiles = p.tiles_x * p.tiles_y;
  hipLaunchKernelGGL(k_dof_segments, dim3((unsigned)((tiles + 1 + 255) / 256)), dim3(256), 0, s, p, keys, n, seg);
  hipLaunchKernelGGL(k_dof_fold, dim3((unsigned)tiles), dim3(kTile * kTile), 0, s, p, z, seg, vals, img_in, alpha_in,
                     img_out, alpha_out);
  return hipGetLastError();
}

// CompileImage (src/cimage.cpp:132-213), one lane per pixel, in the order of
// mbr_compile_image: FP64 brightness and contrast (each rounded to FP32),
// clamp, tanh, clamp,
// 16-bit index into the gamma table.
__global__ void __launch_bounds__(256) k_compile(CompileParams p, const float *__restrict__ r,
                                                 const float *__restrict__ g, const float *__restrict__ b,
                                                 const uint16_t *__restrict__ gamma, uint16_t *__restrict__ out) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= p.n) return;
  const float in[3] = {r[i], g[i], b[i]};
  for (int k = 0; k < 3; k++) {
    float v = (float)(in[k] * p.brightness);
    v = (float)((v - 0.5) * p.contrast + 0.5);
    if (v < 0.0f) v = 0.0f;
    if (p.hdr) v = tanhf(v);
    if (v > 1.0f) v = 1.0f;
    out[3 * (size_t)i + k] = gamma[(uint16_t)(v * 65535.0)];
  }
}

hipError_t launch_compile(const CompileParams &p, const float *r, const float *g, const float *b,
                          const uint16_t *gamma, uint16_t *image16, hipStream_t s) {
  hipLaunchKernelGGL(k_compile, dim3((unsigned)((p.n + 255) / 256)), dim3(256), 0, s, p, r, g, b, gamma, image16);
  return hipGetLastError();
}

hipError_t launch_ssao(const SsaoParams &p, const float *z, const uint16_t *opacity, const uint8_t *const colour[3],
                       uint16_t *image16, hipStream_t s) {
  const dim3 grid((unsigned)((p.width + kTile - 1) / kTile), (unsigned)((p.height + kTile - 1) / kTile));
  hipLaunchKernelGGL(k_ssao, grid, dim3(kTile * kTile), 0, s, p, z, opacity, colour[0], colour[1], colour[2],
                     image16);
  return hipGetLastError();
}

}  // namespace mbr
