// mbr_adjoint.hip — gradient of a pixel loss with respect to the scene
// parameters (formula power, camera position and angles, light angles) for
// FP64 Mandelbulb renders with refined hits (mbr_scene.hit_refine = 1).
//
// The caller supplies the adjoint of the image, dL/dC for every pixel and
// channel.  The kernel returns sum_pixels sum_channels dL/dC . dC/dtheta for
// the nine parameters theta = (power, vp x/y/z, alfa, beta, gamma, light
// alfa, light beta): a vector-Jacobian product, formed per pixel from the
// nine tangents of the pixel's colour, then reduced over the image.
//
// What the colour of a pixel depends on, in the order of the forward:
//   camera: R = Rz(alfa) Rx(beta) Ry(gamma) (src/algebra.cpp:94-143), ray
//     start o = vp - R (0, zoom/persp, 0), view vector d = R (x2 fov, 1, z2 fov)
//     (src/Render3D.cpp:316-333, :389-438);
//   hit: the refined hit t solves F(t) = D(o + t d; power) - c0 thr(t) = 0
//     with thr(t) = t res persp / quality and c0 = 1 - 0.005 quality, so by
//     the implicit function theorem dt = -(dD/dtheta at fixed t) / (dD/dt -
//     c0 dthr/dt), and the hit point p = o + t d moves by do + t dd + d dt;
//   normal: forward differences of D at p (N x 5 iterations, minN 0) with the
//     step delta = thr quality smoothness (src/shaders.cpp:675-757), each
//     distance estimate carrying its gradient in (p, power);
//   light L = R (cos(aL - pi/2) cos(-bL), sin(aL - pi/2) cos(-bL), sin(-bL))
//     (src/Render3D.cpp:253-256);
//   shading: main light x direct light x ((1 - shading) + shading max(0,
//     n.L)) x colour + main light x specular min(15, max(0, n.h)^30) with h =
//     normalize(L - d) (ObjectShader, src/shaders.cpp:21-112); shadows, AO,
//     glow and fog are rejected by mbr_adjoint;
//   misses: the background gradient on d.(0,0,-1) (src/shaders.cpp:159-181).
// The surface colour is an integer palette index (floor of the orbit value,
// src/shaders.cpp:855-870) and the background is stored as 16-bit integers:
// both are steps in theta whose derivative is zero almost everywhere; the
// kernel uses zero for the colour and the slope of the gradient before the
// 16-bit rounding for the background.  Pixels whose depth or normal is
// ill-conditioned (grazing hits, degenerate normals) are left out.
#include "mbr_kernels.hpp"
#include "mbr_shade.hpp"

namespace mbr {

constexpr int kAdjBlock = 256;
constexpr int kParams = 9;
constexpr double kTangentMax = 1e5;
// residual of a refined hit, relative to its threshold, below which it
// counts as a root of F (the refinement stops at 1e-13)
constexpr double kRootTol = 1e-7;

// Value and gradient in (p.x, p.y, p.z, power).
struct G4 {
  double v, g[4];
};

__device__ __forceinline__ G4 gc(double v) { return G4{v, {0, 0, 0, 0}}; }
__device__ __forceinline__ G4 operator+(const G4 &a, const G4 &b) {
  return G4{a.v + b.v, {a.g[0] + b.g[0], a.g[1] + b.g[1], a.g[2] + b.g[2], a.g[3] + b.g[3]}};
}
__device__ __forceinline__ G4 operator-(const G4 &a, const G4 &b) {
  return G4{a.v - b.v, {a.g[0] - b.g[0], a.g[1] - b.g[1], a.g[2] - b.g[2], a.g[3] - b.g[3]}};
}
__device__ __forceinline__ G4 operator*(const G4 &a, const G4 &b) {
  G4 r;
  r.v = a.v * b.v;
  for (int i = 0; i < 4; i++) r.g[i] = a.g[i] * b.v + a.v * b.g[i];
  return r;
}
__device__ __forceinline__ G4 operator*(const G4 &a, double s) {
  return G4{a.v * s, {a.g[0] * s, a.g[1] * s, a.g[2] * s, a.g[3] * s}};
}
__device__ __forceinline__ G4 operator/(const G4 &a, const G4 &b) {
  G4 r;
  r.v = a.v / b.v;
  const double ib = 1.0 / b.v;
  for (int i = 0; i < 4; i++) r.g[i] = (a.g[i] - r.v * b.g[i]) * ib;
  return r;
}
// f(a) with f' = fp at a.v
__device__ __forceinline__ G4 chain(const G4 &a, double fv, double fp) {
  return G4{fv, {a.g[0] * fp, a.g[1] * fp, a.g[2] * fp, a.g[3] * fp}};
}
__device__ __forceinline__ G4 gsqrt(const G4 &a) {
  const double s = sqrt(a.v);
  return chain(a, s, s > 0 ? 0.5 / s : 0.0);
}
__device__ __forceinline__ G4 gasin(const G4 &a) { return chain(a, asin(a.v), 1.0 / sqrt(1.0 - a.v * a.v)); }
__device__ __forceinline__ G4 gcos(const G4 &a) { return chain(a, cos(a.v), -sin(a.v)); }
__device__ __forceinline__ G4 gsin(const G4 &a) { return chain(a, sin(a.v), cos(a.v)); }
__device__ __forceinline__ G4 glog(const G4 &a) { return chain(a, log(a.v), 1.0 / a.v); }
__device__ __forceinline__ G4 gatan2(const G4 &y, const G4 &x) {
  G4 r;
  r.v = atan2(y.v, x.v);
  const double den = x.v * x.v + y.v * y.v;
  const double ix = den > 0 ? x.v / den : 0.0, iy = den > 0 ? y.v / den : 0.0;
  for (int i = 0; i < 4; i++) r.g[i] = ix * y.g[i] - iy * x.g[i];
  return r;
}
// a^e for a > 0 with both a and e carrying gradients
__device__ __forceinline__ G4 gpow(const G4 &a, const G4 &e) {
  G4 r;
  r.v = pow(a.v, e.v);
  const double la = log(a.v), da = e.v * r.v / a.v;
  for (int i = 0; i < 4; i++) r.g[i] = da * a.g[i] + r.v * la * e.g[i];
  return r;
}

struct G4V {
  G4 x, y, z;
};

// CalculateDistance of the trig_optim Mandelbulb (src/fractal.cpp:211-225,
// :1151-1157, :1245-1306 for the analytic branch without interior mode,
// iteration threshold, primitives or limits) with its gradient in (point,
// power); the branches follow the values, as in the forward.
__device__ G4 de_grad(const FractalK<double> &f, const V3<double> &p, double power, double detail, int N, int minN) {
  G4V z = {G4{p.x, {1, 0, 0, 0}}, G4{p.y, {0, 1, 0, 0}}, G4{p.z, {0, 0, 1, 0}}};
  const G4 pw = G4{power, {0, 0, 0, 1}};
  const G4 pm1 = pw - gc(1.0);
  G4V c;
  if (f.julia_mode) {
    c = {gc(f.julia[0]), gc(f.julia[1]), gc(f.julia[2])};
  } else {
    c = {z.x * f.constant_factor, z.y * f.constant_factor, z.z * f.constant_factor};
  }
  G4 r = gsqrt(z.x * z.x + z.y * z.y + z.z * z.z);
  G4 aux = gc(1.0);
  G4 dist = gc(0.0);
  int L = 0;
  bool esc = false;
  for (; L < N; L++) {
    const G4 th0 = gasin(z.z / r);
    const G4 ph0 = gatan2(z.y, z.x);
    G4 rp = gpow(r, pm1);
    const G4 th = th0 * pw, ph = ph0 * pw;
    const G4 cth = gcos(th);
    aux = rp * aux * pw + gc(1.0);
    rp = rp * r;
    z = {cth * gcos(ph) * rp + c.x, cth * gsin(ph) * rp + c.y, gsin(th) * rp + c.z};
    r = gsqrt(z.x * z.x + z.y * z.y + z.z * z.z);
    if (!(r.v <= f.bailout)) {
      dist = r * glog(r) * 0.5 / aux;
      esc = true;
      break;
    }
  }
  if (!esc) dist = gc(0.0);
  if (dist.v < 0) dist = gc(0.0);
  if (L < minN && dist.v < detail) dist = gc(detail);
  return dist;
}

// Rotation Rz(a) Rx(b) Ry(g) and its derivatives in a, b, g (row-major).
__device__ void rot3(double a, double b, double g, double R[9], double dR[3][9]) {
  const double ca = cos(a), sa = sin(a), cb = cos(b), sb = sin(b), cg = cos(g), sg = sin(g);
  const double Z[9] = {ca, -sa, 0, sa, ca, 0, 0, 0, 1}, dZ[9] = {-sa, -ca, 0, ca, -sa, 0, 0, 0, 0};
  const double X[9] = {1, 0, 0, 0, cb, -sb, 0, sb, cb}, dX[9] = {0, 0, 0, 0, -sb, -cb, 0, cb, -sb};
  const double Y[9] = {cg, 0, sg, 0, 1, 0, -sg, 0, cg}, dY[9] = {-sg, 0, cg, 0, 0, 0, -cg, 0, -sg};
  auto mul3 = [](const double *A, const double *B, const double *C, double *out) {
    double t[9];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) t[i * 3 + j] = A[i * 3] * B[j] + A[i * 3 + 1] * B[3 + j] + A[i * 3 + 2] * B[6 + j];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) out[i * 3 + j] = t[i * 3] * C[j] + t[i * 3 + 1] * C[3 + j] + t[i * 3 + 2] * C[6 + j];
  };
  mul3(Z, X, Y, R);
  mul3(dZ, X, Y, dR[0]);
  mul3(Z, dX, Y, dR[1]);
  mul3(Z, X, dY, dR[2]);
}

__device__ __forceinline__ V3<double> mv(const double *M, const V3<double> &v) {
  return v3<double>(M[0] * v.x + M[1] * v.y + M[2] * v.z, M[3] * v.x + M[4] * v.y + M[5] * v.z,
                    M[6] * v.x + M[7] * v.y + M[8] * v.z);
}

// d/dtheta of normalize(v), given v and its tangent dv
__device__ __forceinline__ V3<double> dnormalize(const V3<double> &v, const V3<double> &dv) {
  const double l = length(v);
  const V3<double> u = v * (1.0 / l);
  return (dv - u * dot(u, dv)) * (1.0 / l);
}

template <class RT>
__global__ void __launch_bounds__(kAdjBlock) k_adjoint(const DevScene *__restrict__ scp, BandBufs b, int band_first,
                                                       int band_rays, AdjParams P, const float *__restrict__ gR,
                                                       const float *__restrict__ gG, const float *__restrict__ gB,
                                                       double *grad, uint8_t *skipped, float *__restrict__ jac) {
  const DevScene &sc = *scp;
  const ShadeK<double> &k = sc.sd;
  const FractalK<double> &f = sc.fd;
  double acc[kParams];
  for (int i = 0; i < kParams; i++) acc[i] = 0;
  double R[9], dR[3][9];
  rot3(P.alpha, P.beta, P.gamma, R, dR);
  // light: L = R lv(aL, bL)
  const V3<double> lv = v3<double>(cos(P.light_alpha - 0.5 * M_PI) * cos(-P.light_beta),
                                   sin(P.light_alpha - 0.5 * M_PI) * cos(-P.light_beta), sin(-P.light_beta));
  const V3<double> dlv_a = v3<double>(-sin(P.light_alpha - 0.5 * M_PI) * cos(-P.light_beta),
                                      cos(P.light_alpha - 0.5 * M_PI) * cos(-P.light_beta), 0.0);
  const V3<double> dlv_b = v3<double>(cos(P.light_alpha - 0.5 * M_PI) * sin(-P.light_beta),
                                      sin(P.light_alpha - 0.5 * M_PI) * sin(-P.light_beta), -cos(-P.light_beta));
  const V3<double> L = mv(R, lv);
  V3<double> dL[kParams];
  for (int i = 0; i < kParams; i++) dL[i] = v3<double>(0, 0, 0);
  for (int a = 0; a < 3; a++) dL[4 + a] = mv(dR[a], lv);
  dL[7] = mv(R, dlv_a);
  dL[8] = mv(R, dlv_b);
  const V3<double> yz = v3<double>(0.0, P.zoom_persp, 0.0);  // start = vp - R (0, zoom/persp, 0)
  const double kthr = k.resolution * k.persp / k.quality;
  const double c0 = 1.0 - 0.005 * k.quality;
  const double mlR = k.main_light[0], mlG = k.main_light[1], mlB = k.main_light[2];

  for (long long rr = (long long)blockIdx.x * kAdjBlock + threadIdx.x; rr < band_rays;
       rr += (long long)gridDim.x * kAdjBlock) {
    const int r = (int)rr;
    const int p = band_first + r;
    const int x = p % sc.width, y = p / sc.width;
    const size_t npix = (size_t)sc.width * sc.height;
    const double aR = jac ? 0.0 : gR[p], aG = jac ? 0.0 : gG[p], aB = jac ? 0.0 : gB[p];
    if (!jac && aR == 0.0 && aG == 0.0 && aB == 0.0) continue;
    if (jac)
      for (int i = 0; i < 3 * kParams; i++) jac[(size_t)i * npix + p] = 0.0f;
    const double x2 = ((double)x / sc.width - 0.5) * sc.aspect;
    const double z2 = (double)y / sc.height - 0.5;
    const V3<double> v0 = v3<double>(x2 * sc.fov, 1.0, z2 * sc.fov);
    V3<double> d;  // the forward's view vector (its values); tangents from rot3
    view_vector(sc, x, y, d);
    V3<double> dd[kParams], dO[kParams];
    for (int i = 0; i < kParams; i++) {
      dd[i] = v3<double>(0, 0, 0);
      dO[i] = v3<double>(0, 0, 0);
    }
    for (int a = 0; a < 3; a++) {
      dd[4 + a] = mv(dR[a], v0);
      dO[4 + a] = mv(dR[a], yz) * -1.0;
    }
    if (skipped) skipped[p] = 0;
    dO[1] = v3<double>(1, 0, 0);
    dO[2] = v3<double>(0, 1, 0);
    dO[3] = v3<double>(0, 0, 1);
    double dC[3][kParams];
    if (!b.found[r]) {
      // background gradient on the view vector (src/shaders.cpp:159-181)
      const double grad = -d.z + 1.0;
      for (int i = 0; i < kParams; i++) {
        const double dg = -dd[i].z;
        for (int ch = 0; ch < 3; ch++) {
          const double c1 = sc.bg1[ch], c2 = sc.bg2[ch], c3 = sc.bg3[ch];
          const double slope = grad < 1 ? (c2 - c3) : (c1 - c2);
          dC[ch][i] = slope * dg / 65536.0;
        }
      }
    } else {
      const double t = b.depth[r];
      const V3<double> ph = v3<double>(b.px[r], b.py[r], b.pz[r]);  // the forward's hit point
      const double thr = b.thr[r];
      // the implicit hit
      const G4 D = de_grad(f, ph, P.power, thr, sc.N, sc.minN);
      const V3<double> gD = v3<double>(D.g[0], D.g[1], D.g[2]);
      const double Ft = dot(gD, d) - c0 * kthr;
      // The depth is well conditioned where the ray meets the level set
      // D = c0 thr going down it: Ft < 0 with a margin.  Near a grazing
      // tangency Ft -> 0 and dt/dtheta diverges: the refined hit then sits
      // next to a fold of t(theta), where the forward jumps for a small
      // change of theta; such pixels are left out (and flagged).
      // The implicit-function derivative holds at a root of F.  Where the
      // distance estimate jumps across the binary search's band (its
      // spherical coordinates are singular on the z axis and its preimages)
      // the refinement stops short of a root, and the depth then follows
      // the Newton iterates instead: such hits are left out (and flagged).
      const bool at_root = fabs(D.v - c0 * thr) <= kRootTol * thr;
      if (!(Ft < -0.05 * length(gD) * length(d)) || !at_root) {
        if (skipped) skipped[p] = 1;
        continue;
      }
      if (skipped) skipped[p] = 0;
      double dt[kParams];
      V3<double> dp[kParams];
      for (int i = 0; i < kParams; i++) {
        const double Fi = dot(gD, dO[i] + dd[i] * t) + (i == 0 ? D.g[3] : 0.0);
        dt[i] = fabs(Ft) > 0 ? -Fi / Ft : 0.0;
        dp[i] = dO[i] + dd[i] * t + d * dt[i];
      }
      // normal: forward differences of D at N x 5, minN 0
      const double delta = thr * k.quality * k.smoothness;
      const int N5 = sc.N * 5;
      const G4 s1 = de_grad(f, ph, P.power, thr, N5, 0);
      const G4 s2 = de_grad(f, ph + v3<double>(delta, 0, 0), P.power, thr, N5, 0);
      const G4 s3 = de_grad(f, ph + v3<double>(0, delta, 0), P.power, thr, N5, 0);
      const G4 s4 = de_grad(f, ph + v3<double>(0, 0, delta), P.power, thr, N5, 0);
      const V3<double> nr = v3<double>(s2.v - s1.v, s3.v - s1.v, s4.v - s1.v);
      const bool zero_n = nr.x == 0 && nr.y == 0 && nr.z == 0;
      const V3<double> n = zero_n ? v3<double>(1, 0, 0) : normalize(nr);
      // light and specular (ObjectShader without shadow and ambient terms)
      const double nl = dot(n, L);
      const double shade = k.direct_light * ((1.0 - k.shading) + k.shading * (nl > 0 ? nl : 0.0));
      const V3<double> hr = L - d;
      const V3<double> h = normalize(hr);
      const double nh = dot(n, h);
      const double s2v = nh > 0 ? pow(nh, 30.0) : 0.0;
      const bool spec_live = nh > 0 && s2v < 15.0;
      // colour: the palette index is an integer (derivative zero); the
      // forward's colour, recomputed at the same point
      ShaderIn<double> in;
      in.point = ph;
      in.point64 = ph;
      in.detail = thr;
      in.object = b.object[r];
      LaneCount cnt = {0, 0};
      const RGBT<double> colour = surface_colour<double>(sc, in, cnt);
      const double col[3] = {colour.R, colour.G, colour.B};
      for (int i = 0; i < kParams; i++) {
        const double ddelta = k.quality * k.smoothness * kthr * dt[i];
        const V3<double> q = dp[i];
        const double pw = i == 0 ? 1.0 : 0.0;
        const double ds1 = s1.g[0] * q.x + s1.g[1] * q.y + s1.g[2] * q.z + s1.g[3] * pw;
        const double ds2 = s2.g[0] * (q.x + ddelta) + s2.g[1] * q.y + s2.g[2] * q.z + s2.g[3] * pw;
        const double ds3 = s3.g[0] * q.x + s3.g[1] * (q.y + ddelta) + s3.g[2] * q.z + s3.g[3] * pw;
        const double ds4 = s4.g[0] * q.x + s4.g[1] * q.y + s4.g[2] * (q.z + ddelta) + s4.g[3] * pw;
        const V3<double> dnr = v3<double>(ds2 - ds1, ds3 - ds1, ds4 - ds1);
        const V3<double> dn = zero_n ? v3<double>(0, 0, 0) : dnormalize(nr, dnr);
        const double dnl = dot(dn, L) + dot(n, dL[i]);
        const double dshade = nl > 0 ? k.direct_light * k.shading * dnl : 0.0;
        const V3<double> dh = dnormalize(hr, dL[i] - dd[i]);
        const double dnh = dot(dn, h) + dot(n, dh);
        const double dspec = spec_live ? 30.0 * pow(nh, 29.0) * dnh * k.specular : 0.0;
        dC[0][i] = mlR * dshade * col[0] + mlR * dspec;
        dC[1][i] = mlG * dshade * col[1] + mlG * dspec;
        dC[2][i] = mlB * dshade * col[2] + mlB * dspec;
      }
    }
    // A tangent beyond kTangentMax colour units per unit parameter comes from
    // a degenerate normal (the four estimates nearly equal) or a fold of the
    // hit: the pixel is ill-conditioned and left out (and flagged).
    bool finite = true;
    for (int i = 0; i < kParams; i++)
      for (int ch = 0; ch < 3; ch++) finite = finite && fabs(dC[ch][i]) <= kTangentMax;
    if (!finite) {
      if (skipped) skipped[p] = 1;
      continue;
    }
    if (jac) {
      for (int i = 0; i < kParams; i++)
        for (int ch = 0; ch < 3; ch++) jac[(size_t)(i * 3 + ch) * npix + p] = (float)dC[ch][i];
      continue;
    }
    for (int i = 0; i < kParams; i++) acc[i] += aR * dC[0][i] + aG * dC[1][i] + aB * dC[2][i];
  }
  // block reduction, one atomic per parameter and workgroup
  __shared__ double red[kParams][kAdjBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int i = 0; i < kParams; i++) {
    double v = acc[i];
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    if (lane == 0) red[i][wave] = v;
  }
  __syncthreads();
  if (grad && threadIdx.x < kParams) {
    double v = 0;
    for (int w = 0; w < kAdjBlock / 64; w++) v += red[threadIdx.x][w];
    atomicAdd(&grad[threadIdx.x], v);
  }
}

hipError_t launch_adjoint(const DevScene *sc, const BandBufs &b, int band_first, int band_rays, AdjParams p,
                          const float *gR, const float *gG, const float *gB, double *grad, uint8_t *skipped,
                          float *jac, int blocks, hipStream_t s) {
  if (band_rays <= 0) return hipSuccess;
  if (b.rec_bytes == 8)
    hipLaunchKernelGGL((k_adjoint<double>), dim3(std::max(1, blocks)), dim3(kAdjBlock), 0, s, sc, b, band_first,
                       band_rays, p, gR, gG, gB, grad, skipped, jac);
  else
    hipLaunchKernelGGL((k_adjoint<float>), dim3(std::max(1, blocks)), dim3(kAdjBlock), 0, s, sc, b, band_first,
                       band_rays, p, gR, gG, gB, grad, skipped, jac);
  return hipGetLastError();
}

}  // namespace mbr
