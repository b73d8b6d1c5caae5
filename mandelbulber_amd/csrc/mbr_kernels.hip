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
// ObjectShader (src/shaders.cpp:21-112) runs as three launches:
//   k_surface  one lane per hit: normals, shadow, specular, palette colour
//              and, in fast mode, FastAmbientOcclusion2;
//   k_ao       persistent: the AmbientOcclusion rays of all hits;
//   k_compose  one lane per hit: AO sum in direction order and the final
//              ObjectShader composition.
template <class T>
struct SurfBufs {
  T *base;    // [3][n] mainLight * shade * shadow
  T *spec;    // [3][n] mainLight * specular * shadow
  T *col;     // [3][n] surface colour
  T *fast_ao; // [n]
  T *ao;      // [n * dirs] AO intensity per (hit, direction)
  long long n;
};

template <class T>
__device__ __forceinline__ void load_hit(const DevScene &sc, const BandBufs &b, int r, ShaderIn<T> &in) {
  const ShadeK<T> &k = sc.shade<T>();
  in.point = v3<T>((T)b.px[r], (T)b.py[r], (T)b.pz[r]);
  in.light = v3<T>(k.light[0], k.light[1], k.light[2]);
  in.dist_thresh = (T)b.thr[r];
  in.lastDist = (T)b.last[r];
  in.depth = (T)b.depth[r];
  in.delta = sc.constant_thresh ? (T)(b.depth[r] * sc.sd.resolution * sc.sd.persp) : (T)(b.thr[r] * sc.sd.quality);
  in.detail = in.dist_thresh;
  in.object = b.object[r];
}

template <class T>
__global__ void __launch_bounds__(kBlock) k_surface(const DevScene *__restrict__ scp, BandBufs b, int band_first,
                                                    int n_hits, SurfBufs<T> sb, unsigned long long *stats) {
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
    load_hit(sc, b, r, in);
    in.view = cvt<T>(dir);
    in.normal = normals<T>(sc, in, cnt);
    T sh = dot(in.normal, in.light);
    if (sh < 0) sh = 0;
    const T shade = k.direct_light * ((T(1) - k.shading) + k.shading * sh);
    T shadow = 1;
    if (sc.shadow) shadow = main_shadow<T>(sc, in, cnt);
    V3<T> half = normalize(in.light - in.view);
    T shade2 = dot(in.normal, half);
    if (shade2 < T(0)) shade2 = 0;
    shade2 = Math<T>::pow(shade2, T(30));
    if (shade2 > T(15)) shade2 = 15;
    const T specular = shade2 * k.specular;
    RGBT<T> colour = surface_colour<T>(sc, in, cnt);
    const long long n = sb.n;
    for (int c = 0; c < 3; c++) {
      sb.base[c * n + tid] = k.main_light[c] * shade * shadow;
      sb.spec[c * n + tid] = k.main_light[c] * specular * shadow;
    }
    sb.col[0 * n + tid] = colour.R;
    sb.col[1 * n + tid] = colour.G;
    sb.col[2 * n + tid] = colour.B;
    if (sc.ao && sc.fast_ao) sb.fast_ao[tid] = fast_ao2<T>(sc, in, cnt);
    b.col_r[r] = x86_d2u8((double)colour.R * 255);
    b.col_g[r] = x86_d2u8((double)colour.G * 255);
    b.col_b[r] = x86_d2u8((double)colour.B * 255);
  }
  unsigned long long v[ST_COUNT] = {0};
  v[ST_SHADE_DE] = cnt.de_calls;
  v[ST_SHADE_ITERS] = cnt.iterations;
  flush_stats(stats, v);
}

// AmbientOcclusion (src/shaders.cpp:607-673) for all (hit, direction) pairs
// of a band.  Each lane holds one AO ray and advances it one formula
// iteration per loop pass: when its distance estimate completes it takes the
// ray's next step, and when the ray ends it writes the ray's intensity and
// takes the next (hit, direction) item.  Items are handed out per wave from
// a global counter in chunks, so every lane of every wave does formula work
// until the band's AO rays run out, whatever the spread of per-ray step and
// iteration counts.
constexpr unsigned kAOChunk = 256;

template <class T, int F>
__global__ void __launch_bounds__(kBlock) k_ao(const DevScene *__restrict__ scp, BandBufs b, int n_hits, AODev ao,
                                               SurfBufs<T> sb, unsigned long long *queue,
                                               unsigned long long *stats) {
  const DevScene &sc = *scp;
  const FractalK<T> &f = sc.frac<T>();
  const ShadeK<T> &k = sc.shade<T>();
  const T *dirs;
  if constexpr (sizeof(T) == 8) dirs = ao.dir_d; else dirs = ao.dir_f;
  const unsigned count = (unsigned)ao.count;
  const unsigned long long total = (unsigned long long)n_hits * count;
  const int N = sc.N, minN = sc.minN;
  const int lane = __lane_id();
  const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

  unsigned long long wnext = 0, wend = 0;  // wave-uniform chunk cursor
  bool active = false, exhausted = false;
  unsigned long long item = 0;
  V3<T> P = v3<T>(0, 0, 0), dir = P;
  T r = 0, end_dist = 1, detail = 0, thr_f = 0;
  int guard = 0;
  DECall<T> call;
  LaneCount cnt = {0, 0};

  for (;;) {
    const bool need = !active && !exhausted;
    const unsigned long long mask = __ballot(need);
    if (mask) {
      // hand out the next items to the lanes that need one, in lane order
      const unsigned n = (unsigned)__popcll(mask);
      const unsigned off = (unsigned)__popcll(mask & lt_mask);
      const unsigned long long avail = wend - wnext;
      unsigned long long mine = wnext + off;
      if (n > avail) {
        unsigned long long cbase = 0;
        if (lane == 0) cbase = atomicAdd(queue, (unsigned long long)kAOChunk);
        cbase = __shfl(cbase, 0);
        if (off >= avail) mine = cbase + (off - avail);
        wnext = cbase + (n - avail);
        wend = cbase + kAOChunk;
      } else {
        wnext += n;
      }
      if (need) {
        if (mine >= total) {
          exhausted = true;
        } else {
          item = mine;
          const unsigned h = (unsigned)(item / count), d = (unsigned)(item - (unsigned long long)h * count);
          const int ray = b.hit_list[h];
          ShaderIn<T> in;
          load_hit(sc, b, ray, in);
          P = in.point;
          dir = v3<T>(dirs[4 * d + 0], dirs[4 * d + 1], dirs[4 * d + 2]);
          r = in.delta;                      // start_dist
          end_dist = in.delta / k.resolution;
          detail = in.detail;
          thr_f = (T)(float)in.dist_thresh;  // float dist_thresh (src/shaders.cpp:642)
          guard = 0;
          if (r < end_dist) {
            DEInfo info;
            T dist;
            if (de_call_begin<T, F>(f, P + dir * r, detail, N, minN, false, call, dist, info)) {
              // finished without iterating: treat as the step's result below
              active = true;
              call.it.L = -1;  // marks a completed call
              call.it.r = dist;
              call.it.aux = (T)info.max_iter;
            } else {
              active = true;
            }
          } else {
            sb.ao[item] = T(1);
          }
        }
      }
    }
    if (!__any(active || !exhausted)) break;
    if (active) {
      T dist = 0;
      DEInfo info;
      bool done;
      int mi;
      if (call.it.L < 0) {
        done = true;
        dist = call.it.r;
        mi = call.it.aux != T(0);
        info.work = 0;
      } else {
        done = de_call_step<T, F>(f, call, detail, N, minN, false, dist, info);
        mi = info.max_iter;
      }
      if (done) {
        cnt.de_calls += 1;
        cnt.iterations += (uint32_t)info.work;
        if (dist < thr_f || mi) {
          T shadowTemp = T(1) - (end_dist - r) / end_dist;
          if (shadowTemp < T(0)) shadowTemp = 0;
          sb.ao[item] = shadowTemp;
          active = false;
        } else {
          r += dist * T(2);
          guard++;
          if (!(r < end_dist) || guard >= kMaxShadeSteps) {
            sb.ao[item] = T(1);
            active = false;
          } else {
            T d2;
            DEInfo i2;
            if (de_call_begin<T, F>(f, P + dir * r, detail, N, minN, false, call, d2, i2)) {
              call.it.L = -1;
              call.it.r = d2;
              call.it.aux = (T)i2.max_iter;
            }
          }
        }
      }
    }
  }
  unsigned long long v[ST_COUNT] = {0};
  v[ST_SHADE_DE] = cnt.de_calls;
  v[ST_SHADE_ITERS] = cnt.iterations;
  flush_stats(stats, v);
}

template <class T, class RT>
__global__ void __launch_bounds__(kBlock) k_compose(const DevScene *__restrict__ scp, BandBufs b, int n_hits, AODev ao,
                                                    SurfBufs<T> sb) {
  const DevScene &sc = *scp;
  const int tid = blockIdx.x * kBlock + threadIdx.x;
  if (tid >= n_hits) return;
  const ShadeK<T> &k = sc.shade<T>();
  const int r = b.hit_list[tid];
  const long long n = sb.n;
  RGBT<T> amb = {0, 0, 0};
  if (sc.ao) {
    if (sc.fast_ao) {
      const T a = sb.fast_ao[tid];
      amb = RGBT<T>{a, a, a};
    } else {
      const T *it = sb.ao + (long long)tid * ao.count;
      for (int v = 0; v < ao.count; v++) {
        const T s = it[v];
        amb.R += s * T(ao.rgb[4 * v + 0]);
        amb.G += s * T(ao.rgb[4 * v + 1]);
        amb.B += s * T(ao.rgb[4 * v + 2]);
      }
      const T den = T(ao.count) * T(256);
      amb.R /= den;
      amb.G /= den;
      amb.B /= den;
    }
  }
  const T cR = sb.col[0 * n + tid], cG = sb.col[1 * n + tid], cB = sb.col[2 * n + tid];
  ((RT *)b.sh_r)[r] = (RT)((amb.R * k.global_ilum + sb.base[0 * n + tid]) * cR + sb.spec[0 * n + tid]);
  ((RT *)b.sh_g)[r] = (RT)((amb.G * k.global_ilum + sb.base[1 * n + tid]) * cG + sb.spec[1 * n + tid]);
  ((RT *)b.sh_b)[r] = (RT)((amb.B * k.global_ilum + sb.base[2 * n + tid]) * cB + sb.spec[2 * n + tid]);
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
                        ShadeBufs sbuf, int cu_count, unsigned long long *stats, hipStream_t s) {
  if (n_hits <= 0) return hipSuccess;
  const int formula = sbuf.formula;
  // persistent AO grid: enough workgroups to fill every CU several times
  const dim3 ao_grid((unsigned)(cu_count > 0 ? cu_count * 8 : 2048));
  auto run = [&](auto tag) -> hipError_t {
    using T = decltype(tag);
    SurfBufs<T> sb;
    sb.base = (T *)sbuf.base;
    sb.spec = (T *)sbuf.spec;
    sb.col = (T *)sbuf.col;
    sb.fast_ao = (T *)sbuf.fast_ao;
    sb.ao = (T *)sbuf.ao;
    sb.n = sbuf.n;
    hipLaunchKernelGGL((k_surface<T>), grid_for(n_hits), dim3(kBlock), 0, s, sc, b, band_first, n_hits, sb, stats);
    if (sbuf.slow_ao) {
      hipError_t e = hipMemsetAsync(sbuf.queue, 0, sizeof(unsigned long long), s);
      if (e != hipSuccess) return e;
      if (formula == 7)
        hipLaunchKernelGGL((k_ao<T, 7>), ao_grid, dim3(kBlock), 0, s, sc, b, n_hits, ao, sb, sbuf.queue, stats);
      else if (formula == 8)
        hipLaunchKernelGGL((k_ao<T, 8>), ao_grid, dim3(kBlock), 0, s, sc, b, n_hits, ao, sb, sbuf.queue, stats);
      else
        hipLaunchKernelGGL((k_ao<T, 2>), ao_grid, dim3(kBlock), 0, s, sc, b, n_hits, ao, sb, sbuf.queue, stats);
    }
    hipLaunchKernelGGL((k_compose<T, T>), grid_for(n_hits), dim3(kBlock), 0, s, sc, b, n_hits, ao, sb);
    return hipGetLastError();
  };
  if (precision == 0) return run(double{});
  return run(float{});
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
