// mbr_post.hip — post effects on a rendered frame: screen-space ambient
// occlusion (PostRendering_SSAO / ThreadSSAO, src/image.cpp:249-415).
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

}  // namespace

hipError_t launch_ssao(const SsaoParams &p, const float *z, const uint16_t *opacity, const uint8_t *const colour[3],
                       uint16_t *image16, hipStream_t s) {
  const dim3 grid((unsigned)((p.width + kTile - 1) / kTile), (unsigned)((p.height + kTile - 1) / kTile));
  hipLaunchKernelGGL(k_ssao, grid, dim3(kTile * kTile), 0, s, p, z, opacity, colour[0], colour[1], colour[2],
                     image16);
  return hipGetLastError();
}

}  // namespace mbr
