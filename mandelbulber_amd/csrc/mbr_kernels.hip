// mbr_kernels.hip — gfx950 kernels of the render core.
//
// One frame (or band of rows) runs as three launches:
//   k_march   one lane per primary ray: DE-stepped march with step
//             records, then the binary-search refinement in FP64
//             (RayMarching, src/Render3D.cpp:79-199); hits are appended to a
//             compacted hit list so shading runs on full waves of hits.
//   k_shade   one lane per hit: normals, hard/soft shadow, specular, palette
//             colour, ambient occlusion (ObjectShader, src/shaders.cpp:21).
//   k_finish  one lane per ray: background for misses, VolumetricShader
//             (glow, fog, volumetric fog) over the step record read back to
//             front, and coalesced writes of the framebuffer planes
//             (src/Render3D.cpp:530-589).
#include "mbr_kernels.hpp"
#include "mbr_shade.hpp"

namespace mbr {

constexpr int kBlock = 256;
constexpr double kPi = 3.14159265358979323846;

// per-pixel view vector (src/Render3D.cpp:389-438); true when the fish-eye
// cut removes the pixel
__device__ __forceinline__ bool view_vector(const DevScene &sc, int x, int y, V3<double> &out) {
  const double tiles = sc.tiles;
  const double x2 = ((double)x / sc.width / tiles - 0.5 + sc.tile_x / tiles) * sc.aspect;
  const double z2 = ((double)y / sc.height / tiles - 0.5 + sc.tile_y / tiles);
  const double fov = sc.fov;
  bool cut = false;
  V3<double> v;
  if (sc.persp_type == 1) {
    if (sc.fish_eye_cut && sqrt(x2 * x2 + z2 * z2) > 0.5 / fov) cut = true;
    const double x3 = x2 * kPi, z3 = z2 * kPi;
    const double r = sqrt(x3 * x3 + z3 * z3);
    if (r == 0) v = v3<double>(0.0, 1.0, 0.0);
    else v = v3<double>(x3 / r * sin(r * fov), cos(r * fov), z3 / r * sin(r * fov));
    v = normalize(v);
  } else if (sc.persp_type == 2) {
    const double x3 = x2 * kPi, z3 = z2 * kPi;
    v = v3<double>(sin(fov * x3) * cos(fov * z3), cos(fov * x3) * cos(fov * z3), sin(fov * z3));
    v = normalize(v);
  } else {
    v = v3<double>(x2 * fov, 1.0, z2 * fov);
  }
  out = sc.R_identity ? v : mat_apply(sc.R, v);
  return cut;
}

template <class RT>
struct RecPlanes {
  RT *step, *dist, *thr;
  long long base;   // element offset of record 0
  long long stride; // element distance between records i and i+1
  int limit;        // records beyond this index are not stored
  __device__ __forceinline__ void put(int i, double step_v, double dist_v, double thr_v) const {
    if (i < limit) {
      const long long o = base + (long long)i * stride;
      step[o] = (RT)step_v;
      dist[o] = (RT)dist_v;
      thr[o] = (RT)thr_v;
    }
  }
};

// block-level reduction of per-lane counters into the device stats: LDS
// atomics inside the workgroup, then one global atomic per counter
__device__ __forceinline__ void flush_stats(unsigned long long *stats, const unsigned long long (&v)[ST_COUNT]) {
  __shared__ unsigned long long acc[ST_COUNT];
  if (threadIdx.x < ST_COUNT) acc[threadIdx.x] = 0;
  __syncthreads();
#pragma unroll
  for (int i = 0; i < ST_COUNT; i++)
    if (v[i]) atomicAdd(&acc[i], v[i]);
  __syncthreads();
  if (threadIdx.x < ST_COUNT && acc[threadIdx.x]) atomicAdd(&stats[threadIdx.x], acc[threadIdx.x]);
}

// ----------------------------------------------------------------- march
struct MarchOut {
  V3<double> point;
  double scan, distThresh, dist;
  int buffCount, object;
  bool found, missed;
};

// RayMarching (src/Render3D.cpp:79-199) for one ray.  T is the arithmetic
// of the DE during the march; scan, ray points and thresholds are FP64, and
// the binary search evaluates the DE in FP64.
//
// For T = float the march also bounds its own drift: every FP32 distance
// differs from the FP64 one by at most a few ulp of the distance plus the
// rounding of the point to FP32, and the step taken from it moves the scan
// by that much.  A hit test or binary-search decision whose FP64 margin to
// its threshold is inside twice the accumulated bound could have gone the
// other way in the FP64 march; the function then reports `ambiguous` and the
// caller re-marches the ray in FP64.
template <class T, class RT>
__device__ __forceinline__ void march_ray(const DevScene &sc, V3<double> start, V3<double> dir, uint32_t pixel,
                                          const RecPlanes<RT> &rec, MarchOut &o, LaneCount &cnt, bool &ambiguous) {
  constexpr bool kF32 = sizeof(T) < 8;
  const FractalK<T> &f = sc.frac<T>();
  const ShadeK<double> &k = sc.sd;
  const double quality = k.quality, res = k.resolution, persp = k.persp, DEf = k.DE_factor;
  const double search_limit = 1.0 - 0.01 * quality;
  const double maxScan = sc.view_distance_max;
  const bool fisheye = sc.persp_type == 1 || sc.persp_type == 2;
  const double stepFactor = sc.interior ? 0.8 : 0.5;
  double scan = sc.view_distance_min, distThresh = 0.0, dist = 0.0, step = sc.view_distance_min;
  double detail = 0.0;
  double drift = 0.0;  // bound on |scan_fp32 - scan_fp64|
  int buffCount = 0, object = OBJ_FRACTAL;
  bool found = false, missed = false;
  V3<double> point = v3<double>(0, 0, 0);
  DEInfo info;
  ambiguous = false;
  for (int i = 0; i < 10000; i++) {
    point = start + dir * scan;
    distThresh = sc.constant_thresh ? quality : scan * res * persp / quality + 0.0;
    if (fisheye) distThresh *= kPi;
    detail = distThresh;
    dist = (double)calc_distance<T>(f, cvt<T>(point), (T)detail, sc.N, sc.minN, false, info);
    cnt.add(info);
    double err = 0.0;
    if constexpr (kF32) {
      // FP32 error of this distance: 4 ulp of the distance + the FP32
      // rounding of the point (the DE moves at most ~1 per unit of position)
      const double pmax = fmax(fabs(point.x), fmax(fabs(point.y), fabs(point.z)));
      err = 4.8e-7 * fabs(dist) + 1.2e-7 * pmax;
      // tentative hits and near-threshold steps are decided in FP64
      if (dist < 1.001 * distThresh || dist - distThresh < 2.0 * (drift + err)) {
        dist = calc_distance<double>(sc.fd, point, detail, sc.N, sc.minN, false, info);
        cnt.add(info);
        err = 0.0;
        if (fabs(dist - distThresh) < 2.0 * drift) ambiguous = true;
      }
    }
    object = info.object;
    const double dist_raw = dist;
    if (dist > 3.0) dist = 3.0;
    if (dist < distThresh) {
      if (dist < 0.1 * distThresh) missed = true;
      found = true;
      break;
    }
    const double jf = sc.jitter_mode ? 1.0 - jitter_draw(sc.jitter_seed, sc.frame_no, pixel, i) / 10000.0 : 1.0;
    const double old_step = step;
    step = (dist - stepFactor * distThresh) * DEf * jf;
    rec.put(i, old_step, dist_raw, distThresh);
    buffCount++;
    scan += step;
    if constexpr (kF32) drift += err * DEf;
    if (scan > maxScan) break;
  }
  if (found) {
    // binary search into [thr*(1-0.01q), thr], FP64 (src/Render3D.cpp:145-183)
    step *= 0.5;
    for (int i = 0; i < 20; i++) {
      if constexpr (kF32) {
        if (fabs(dist - distThresh) < 2.0 * drift || fabs(dist - distThresh * search_limit) < 2.0 * drift)
          ambiguous = true;
      }
      if (dist < distThresh && dist > distThresh * search_limit) break;
      if (dist > distThresh) { scan += step; point = start + dir * scan; }
      else if (dist < distThresh * search_limit) { scan -= step; point = start + dir * scan; }
      distThresh = sc.constant_thresh ? quality : scan * res * persp / quality + 0.0;
      DEInfo bi;
      dist = calc_distance<double>(sc.fd, point, detail, sc.N, sc.minN, false, bi);
      cnt.add(bi);
      object = bi.object;
      step *= 0.5;
    }
  }
  o.point = point;
  o.scan = scan;
  o.distThresh = distThresh;
  o.dist = dist;
  o.buffCount = buffCount;
  o.object = object;
  o.found = found;
  o.missed = missed;
}

// One lane per primary ray (or, in re-march mode, per listed ray).  The
// FP32 instantiation re-marches a ray in FP64 when march_ray reports that
// its discrete decisions are within the FP32 error bound.
template <class T, class RT>
__global__ void __launch_bounds__(kBlock) k_march(const DevScene *__restrict__ scp, BandBufs b, int band_first,
                                                  int band_rays, const int *__restrict__ list, int n_list,
                                                  unsigned long long *stats) {
  const DevScene &sc = *scp;
  const int tid = blockIdx.x * kBlock + threadIdx.x;
  const bool remarch = list != nullptr;
  int r = -1;
  if (remarch) { if (tid < n_list) r = list[tid]; }
  else if (tid < band_rays) r = tid;

  LaneCount cnt = {0, 0};
  int hit = 0, missed = 0, redo = 0;
  if (r >= 0) {
    const int p = band_first + r;
    const int x = p % sc.width, y = p / sc.width;
    V3<double> dir;
    const bool cut = view_vector(sc, x, y, dir);
    const V3<double> start = v3<double>(sc.start[0], sc.start[1], sc.start[2]);
    const uint32_t pixel = sc.pixel_base + (uint32_t)p;

    RecPlanes<RT> rec;
    rec.step = (RT *)(remarch ? b.ovf_step : b.rec_step);
    rec.dist = (RT *)(remarch ? b.ovf_dist : b.rec_dist);
    rec.thr = (RT *)(remarch ? b.ovf_thr : b.rec_thr);
    if (remarch) { rec.base = b.ovf_off[r]; rec.stride = 1; rec.limit = b.count[r]; }
    else { rec.base = r; rec.stride = b.cap; rec.limit = b.slab; }

    MarchOut o;
    o.point = v3<double>(0, 0, 0);
    o.scan = 0; o.distThresh = 0; o.dist = 0; o.buffCount = 0; o.object = OBJ_FRACTAL;
    o.found = false; o.missed = false;
    if (!cut) {
      bool ambiguous = false;
      march_ray<T, RT>(sc, start, dir, pixel, rec, o, cnt, ambiguous);
      if constexpr (sizeof(T) < 8) {
        if (ambiguous) {
          bool unused;
          march_ray<double, RT>(sc, start, dir, pixel, rec, o, cnt, unused);
          redo = 1;
        }
      }
    }
    if (!remarch) {
      b.found[r] = o.found;
      b.object[r] = (uint8_t)o.object;
      b.count[r] = o.buffCount;
      b.depth[r] = o.scan;
      b.thr[r] = o.distThresh;
      b.last[r] = o.dist;
      b.px[r] = o.point.x;
      b.py[r] = o.point.y;
      b.pz[r] = o.point.z;
      missed = o.missed;
      if (o.found) {
        const unsigned long long h = atomicAdd(&b.counters[0], 1ull);
        b.hit_list[h] = r;
        hit = 1;
      }
      if (o.buffCount > b.slab) {
        // reserve an exactly sized run of overflow records for the re-march
        const unsigned long long ov = atomicAdd(&b.counters[1], 1ull);
        b.ovf_list[ov] = r;
        b.ovf_off[r] = (long long)atomicAdd(&b.counters[2], (unsigned long long)o.buffCount);
      }
    }
  }
  if (!remarch) {
    unsigned long long v[ST_COUNT] = {0};
    v[ST_MARCH_DE] = cnt.de_calls;
    v[ST_MARCH_ITERS] = cnt.iterations;
    v[ST_HITS] = (unsigned long long)hit;
    v[ST_MISSED] = (unsigned long long)missed;
    v[ST_REDO] = (unsigned long long)redo;
    flush_stats(stats, v);
  }
}

// ----------------------------------------------------------------- shade
template <class T, class RT>
__global__ void __launch_bounds__(kBlock) k_shade(const DevScene *__restrict__ scp, BandBufs b, int band_first,
                                                  int n_hits, AODev ao, unsigned long long *stats) {
  const DevScene &sc = *scp;
  const int tid = blockIdx.x * kBlock + threadIdx.x;
  LaneCount cnt = {0, 0};
  if (tid < n_hits) {
    const int r = b.hit_list[tid];
    const int p = band_first + r;
    const int x = p % sc.width, y = p / sc.width;
    V3<double> dir;
    view_vector(sc, x, y, dir);
    const ShadeK<T> &k = sc.shade<T>();
    ShaderIn<T> in;
    in.point = v3<T>((T)b.px[r], (T)b.py[r], (T)b.pz[r]);
    in.view = cvt<T>(dir);
    in.light = v3<T>(k.light[0], k.light[1], k.light[2]);
    in.dist_thresh = (T)b.thr[r];
    in.lastDist = (T)b.last[r];
    in.depth = (T)b.depth[r];
    in.delta = sc.constant_thresh ? (T)(b.depth[r] * sc.sd.resolution * sc.sd.persp) : (T)(b.thr[r] * sc.sd.quality);
    in.detail = in.dist_thresh;
    in.object = b.object[r];
    AOTable<T> tab;
    if constexpr (sizeof(T) == 8) tab.dir = ao.dir_d; else tab.dir = ao.dir_f;
    tab.rgb = ao.rgb;
    tab.count = ao.count;
    RGBT<T> colour;
    RGBT<T> px = object_shader<T>(sc, in, tab, colour, cnt);
    ((RT *)b.sh_r)[r] = (RT)px.R;
    ((RT *)b.sh_g)[r] = (RT)px.G;
    ((RT *)b.sh_b)[r] = (RT)px.B;
    b.col_r[r] = x86_d2u8((double)colour.R * 255);
    b.col_g[r] = x86_d2u8((double)colour.G * 255);
    b.col_b[r] = x86_d2u8((double)colour.B * 255);
  }
  unsigned long long v[ST_COUNT] = {0};
  v[ST_SHADE_DE] = cnt.de_calls;
  v[ST_SHADE_ITERS] = cnt.iterations;
  flush_stats(stats, v);
}

// ---------------------------------------------------------------- finish
template <class T, class RT>
__global__ void __launch_bounds__(kBlock) k_finish(const DevScene *__restrict__ scp, BandBufs b, int band_first,
                                                   int band_rays, FrameOut out) {
  const DevScene &sc = *scp;
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= band_rays) return;
  const int p = band_first + r;
  const bool found = b.found[r];
  RGBT<T> px;
  T alpha;
  uint8_t cr = 0, cg = 0, cb = 0;
  if (found) {
    px = RGBT<T>{(T)((const RT *)b.sh_r)[r], (T)((const RT *)b.sh_g)[r], (T)((const RT *)b.sh_b)[r]};
    alpha = 1;
    cr = b.col_r[r]; cg = b.col_g[r]; cb = b.col_b[r];
  } else {
    const int x = p % sc.width, y = p / sc.width;
    V3<double> dir;
    view_vector(sc, x, y, dir);
    RGBT<double> bg = background<double>(sc, dir);
    px = RGBT<T>{(T)bg.R, (T)bg.G, (T)bg.B};
    alpha = 0;
  }
  const int n = b.count[r];
  const T depth = (T)b.depth[r];
  VolOut<T> v;
  if (n <= b.slab) {
    const RT *st = (const RT *)b.rec_step, *di = (const RT *)b.rec_dist, *th = (const RT *)b.rec_thr;
    const long long cap = b.cap;
    v = volumetric<T>(sc, depth, n, px, alpha, [&](int i, T &s, T &d, T &t) {
      const long long o = (long long)i * cap + r;
      s = (T)st[o]; d = (T)di[o]; t = (T)th[o];
    });
  } else {
    const RT *st = (const RT *)b.ovf_step, *di = (const RT *)b.ovf_dist, *th = (const RT *)b.ovf_thr;
    const long long base = b.ovf_off[r];
    v = volumetric<T>(sc, depth, n, px, alpha, [&](int i, T &s, T &d, T &t) {
      const long long o = base + i;
      s = (T)st[o]; d = (T)di[o]; t = (T)th[o];
    });
  }
  if (out.rgb[0]) out.rgb[0][p] = (float)v.rgb.R;
  if (out.rgb[1]) out.rgb[1][p] = (float)v.rgb.G;
  if (out.rgb[2]) out.rgb[2][p] = (float)v.rgb.B;
  if (out.z) out.z[p] = found ? (float)b.depth[r] : 1e20f;
  if (out.colour[0]) out.colour[0][p] = cr;
  if (out.colour[1]) out.colour[1][p] = cg;
  if (out.colour[2]) out.colour[2][p] = cb;
  if (out.opacity) out.opacity[p] = x86_d2u16((double)v.opacity * 65535.0);
  if (out.alpha) out.alpha[p] = x86_d2u16((double)v.alpha * 65535.0);
}

// ---------------------------------------------------------------- launch
static inline dim3 grid_for(int n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

hipError_t launch_march(int precision, const DevScene *sc, const BandBufs &b, int band_first, int band_rays,
                        const int *list, int n_list, unsigned long long *stats, hipStream_t s) {
  const int n = list ? n_list : band_rays;
  if (n <= 0) return hipSuccess;
  if (precision == 0)
    hipLaunchKernelGGL((k_march<double, double>), grid_for(n), dim3(kBlock), 0, s, sc, b, band_first, band_rays, list,
                       n_list, stats);
  else
    hipLaunchKernelGGL((k_march<float, float>), grid_for(n), dim3(kBlock), 0, s, sc, b, band_first, band_rays, list,
                       n_list, stats);
  return hipGetLastError();
}

hipError_t launch_shade(int precision, const DevScene *sc, const BandBufs &b, int band_first, int n_hits, AODev ao,
                        unsigned long long *stats, hipStream_t s) {
  if (n_hits <= 0) return hipSuccess;
  if (precision == 0)
    hipLaunchKernelGGL((k_shade<double, double>), grid_for(n_hits), dim3(kBlock), 0, s, sc, b, band_first, n_hits, ao,
                       stats);
  else
    hipLaunchKernelGGL((k_shade<float, float>), grid_for(n_hits), dim3(kBlock), 0, s, sc, b, band_first, n_hits, ao,
                       stats);
  return hipGetLastError();
}

hipError_t launch_finish(int precision, const DevScene *sc, const BandBufs &b, int band_first, int band_rays,
                         FrameOut out, hipStream_t s) {
  if (band_rays <= 0) return hipSuccess;
  if (precision == 0)
    hipLaunchKernelGGL((k_finish<double, double>), grid_for(band_rays), dim3(kBlock), 0, s, sc, b, band_first,
                       band_rays, out);
  else
    hipLaunchKernelGGL((k_finish<float, float>), grid_for(band_rays), dim3(kBlock), 0, s, sc, b, band_first,
                       band_rays, out);
  return hipGetLastError();
}

}  // namespace mbr
