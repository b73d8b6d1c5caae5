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
#include <type_traits>

#include "mbr_kernels.hpp"
#include "mbr_shade.hpp"

namespace mbr {

constexpr int kBlock = 256;
constexpr double kPi = 3.14159265358979323846;


// Step record writer of one ray: records below the slab size go to the
// step-major slab, later ones to chunks of the overflow pool listed in the
// ray's directory; the ray takes its directory (one atomic) on its first
// overflowing record and each chunk on the chunk's first record.  A
// re-march of the same ray reuses them.  When the pool is full the record
// is dropped and ST_POOL_MISS tells the host to grow the pool and render
// the band again.  All records of one ray are written by one lane.
template <class RT>
struct RecPlanes {
  RT *step, *dist, *thr;     // slab planes
  RT *ostep, *odist, *othr;  // overflow pool planes
  int *it, *oit;             // iteration counts (null unless iteration fog is on)
  long long *ovf_off;
  int *dir;
  unsigned long long *dirs, *chunks, *stats;
  long long cap, max_dirs, max_chunks;
  int slab, r;
  __device__ __forceinline__ void put(int i, double step_v, double dist_v, double thr_v, int iters) const {
    long long o;
    RT *st, *di, *th;
    int *itp;
    if (i < slab) {
      o = (long long)i * cap + r;
      st = step; di = dist; th = thr; itp = it;
    } else {
      long long dj = ovf_off[r];
      if (dj == -1) {
        const unsigned long long n = atomicAdd(dirs, 1ull);
        atomicAdd(&stats[ST_OVF_RAYS], 1ull);
        dj = n < (unsigned long long)max_dirs ? (long long)n : -2;
        if (dj >= 0)
          for (int k = 0; k < kDirSlots; k++) dir[dj * kDirSlots + k] = -1;
        ovf_off[r] = dj;
      }
      if (dj < 0) {
        atomicAdd(&stats[ST_POOL_MISS], 1ull);
        return;
      }
      const int rel = i - slab;
      int *slot = &dir[dj * kDirSlots + rel / kChunk];
      int ch = *slot;
      if (ch == -1) {
        const unsigned long long n = atomicAdd(chunks, 1ull);
        atomicAdd(&stats[ST_OVF_CHUNKS], 1ull);
        ch = n < (unsigned long long)max_chunks ? (int)n : -2;
        *slot = ch;
      }
      if (ch < 0) {
        atomicAdd(&stats[ST_POOL_MISS], 1ull);
        return;
      }
      o = (long long)ch * kChunk + rel % kChunk;
      st = ostep; di = odist; th = othr; itp = oit;
    }
    st[o] = (RT)step_v;
    di[o] = (RT)dist_v;
    th[o] = (RT)thr_v;
    if (itp) itp[o] = iters;
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
// RayMarching (src/Render3D.cpp:79-199) is split around its distance
// estimates.  march_before_de sets up step i (point, threshold);
// march_after_de takes step i's distance and either continues, ends the ray
// as a miss, or hands the ray over to the FP64 continuation.  The persistent
// kernel k_march advances many rays this way, one formula iteration per
// pass; the FP64 redo of k_march_fix runs the same functions sequentially,
// so both produce identical records.
//
// Hand-over rule.  The FP64 path hands a ray over when step i is a hit; the
// continuation repeats step i in FP64 and runs the binary search.  The FP32
// path hands a ray over when step i's FP32 distance could be a hit once FP32
// error is allowed for: below 1.001 x threshold, or within the drift margin
// of it.
//
// Drift model.  An FP32 distance differs from the FP64 distance at the same
// point by about 1e-6 of the distance (rounding accumulated over the
// iterations) plus the FP32 rounding of the point, 6e-8 of its largest
// coordinate; the step taken from it (times DE_factor) carries that error
// into the scan.  The errors of successive steps are independent roundings,
// so the scan's deviation from the FP64 march is estimated by their root sum
// of squares, sigma; the margin is 4 sigma.  The continuation decides every
// later hit test and every binary-search branch in FP64.  A decision that
// lies within the margin of its threshold may differ from the FP64 march's.
// Either way the binary search ends in the band where the distance lies in
// (thr (1 - 0.01 quality), thr), next to the point where the decision was
// taken; along the ray the distance changes at rate g (estimated from the
// last two distance estimates), so that band spans about
//   W = (0.01 quality thr + 2 margin) / g
// of depth, and both outcomes end within W of each other.  A branch whose
// step is below 5e-6 of the depth moves the result by less than 2e-5 of the
// depth and is not counted.  The ray is re-marched in FP64 from the start
// when a counted decision was ambiguous and W exceeds kRedoTol of the
// depth, i.e. on grazing incidence.
constexpr double kRedoTol = 3e-5;

struct MarchOut {
  V3<double> point;
  double scan, distThresh, dist;
  int buffCount, object;
  bool found, missed;
};

struct RayState {
  V3<double> point;
  double scan, step, thr, drift, dist;
  double thr0;  // initial threshold: 0, or the previous ray's final one for a reflected ray
  int i, object;
};

enum : int { SO_CONTINUE = 0, SO_MISS = 1, SO_HANDOFF = 2 };

// Refl: a reflected ray, whose thresholds start from s.thr0
// (distThreshInit, src/Render3D.cpp:92, :102-109); primary rays start from 0.
template <bool Refl = false>
__device__ __forceinline__ void march_before_de(const DevScene &sc, V3<double> start, V3<double> dir, RayState &s) {
#pragma clang fp contract(off)
  const ShadeK<double> &k = sc.sd;
  s.point = start + dir * s.scan;
  s.thr = sc.constant_thresh ? k.quality : s.scan * k.resolution * k.persp / k.quality + (Refl ? s.thr0 : 0.0);
  if (sc.persp_type == 1 || sc.persp_type == 2) s.thr *= kPi;
}

// obj_it: the estimate's object type in the low 3 bits, its iteration count
// (CalculateDistance's itersOut) above them.
template <class T, class RT>
__device__ __forceinline__ int march_after_de(const DevScene &sc, RayState &s, double d, int obj_it, uint32_t pixel,
                                              const RecPlanes<RT> &rec) {
#pragma clang fp contract(off)
  const ShadeK<double> &k = sc.sd;
  s.object = obj_it & 7;
  const double dist = d > 3.0 ? 3.0 : d;
  double err = 0.0;
  if constexpr (sizeof(T) < 8) {
    // s.drift holds the variance of the scan deviation
    const double pmax = fmax(fabs(s.point.x), fmax(fabs(s.point.y), fabs(s.point.z)));
    err = 1e-6 * fabs(d) + 6e-8 * pmax;
    if (dist < 1.001 * s.thr || dist - s.thr < 4.0 * sqrt(s.drift + err * err)) return SO_HANDOFF;
  } else {
    if (dist < s.thr) return SO_HANDOFF;
  }
  const double jf = sc.jitter_mode ? 1.0 - jitter_draw(sc.jitter_seed, sc.frame_no, pixel, (uint32_t)s.i) / 10000.0 : 1.0;
  const double old_step = s.step;
  s.step = (dist - (sc.interior ? 0.8 : 0.5) * s.thr) * k.DE_factor * jf;
  rec.put(s.i, old_step, d, s.thr, obj_it >> 3);
  s.scan += s.step;
  s.drift += (err * k.DE_factor) * (err * k.DE_factor);
  s.dist = dist;
  s.i++;
  if (s.scan > sc.view_distance_max || s.i >= 10000) return SO_MISS;
  return SO_CONTINUE;
}

// The reference loop from step s.i on, in FP64, then the binary search
// (src/Render3D.cpp:145-183).  Sets `ambiguous` when a decision lies within
// the drift margin of its threshold (see the hand-over rule above).
template <class RT, int F, bool Refl = false, int IP = 0>
__device__ void march_continue_f64(const DevScene &sc, V3<double> start, V3<double> dir, uint32_t pixel,
                                   const RecPlanes<RT> &rec, RayState s, MarchOut &o, LaneCount &cnt,
                                   bool &ambiguous) {
#pragma clang fp contract(off)
  // trig-free integer-power iteration where the records are FP32, as in
  // k_march
  constexpr bool kAlg = sizeof(RT) < 8 || IP != 0;
  const ShadeK<double> &k = sc.sd;
  const double search_limit = 1.0 - 0.01 * k.quality;
  const double margin = 4.0 * sqrt(s.drift);  // s.drift: variance of the FP32 scan deviation
  bool found = false, missed = false;
  double dist = 0.0, detail = 0.0;
  double prev_dist = s.dist, prev_scan = s.scan - s.step;  // the step before s.i
  bool have_prev = s.i > 0;
  ambiguous = false;
  for (; s.i < 10000; s.i++) {
    march_before_de<Refl>(sc, start, dir, s);
    detail = s.thr;
    DEInfo info;
    dist = calc_distance_f<double, F, kAlg, IP>(sc.fd, s.point, detail, sc.N, sc.minN, false, info);
    cnt.add(info);
    s.object = info.object;
    const double dist_raw = dist;
    if (dist > 3.0) dist = 3.0;
    if (fabs(dist - s.thr) < margin) ambiguous = true;
    if (dist < s.thr) {
      if (dist < 0.1 * s.thr) missed = true;
      found = true;
      break;
    }
    const double jf = sc.jitter_mode ? 1.0 - jitter_draw(sc.jitter_seed, sc.frame_no, pixel, (uint32_t)s.i) / 10000.0 : 1.0;
    const double old_step = s.step;
    s.step = (dist - (sc.interior ? 0.8 : 0.5) * s.thr) * k.DE_factor * jf;
    rec.put(s.i, old_step, dist_raw, s.thr, info.iters);
    prev_dist = dist;
    prev_scan = s.scan;
    have_prev = true;
    s.scan += s.step;
    if (s.scan > sc.view_distance_max) { s.i++; break; }
  }
  double distThresh = s.thr, scan = s.scan;
  V3<double> point = s.point;
  int object = s.object;
  if (found) {
    double step = s.step * 0.5;
    for (int j = 0; j < 20; j++) {
      if (step > 5e-6 * scan &&
          (fabs(dist - distThresh) < margin || fabs(dist - distThresh * search_limit) < margin))
        ambiguous = true;
      if (dist < distThresh && dist > distThresh * search_limit) break;
      prev_dist = dist;
      prev_scan = scan;
      have_prev = true;
      if (dist > distThresh) { scan += step; point = start + dir * scan; }
      else if (dist < distThresh * search_limit) { scan -= step; point = start + dir * scan; }
      distThresh = sc.constant_thresh ? k.quality : scan * k.resolution * k.persp / k.quality + (Refl ? s.thr0 : 0.0);
      DEInfo bi;
      dist = calc_distance_f<double, F, kAlg, IP>(sc.fd, point, detail, sc.N, sc.minN, false, bi);
      cnt.add(bi);
      object = bi.object;
      step *= 0.5;
    }
    if (ambiguous && have_prev) {
      const double g = fabs(dist - prev_dist) / fmax(fabs(scan - prev_scan), 1e-300);
      const double W = (0.01 * k.quality * distThresh + 2.0 * margin) / fmax(g, 1e-12);
      ambiguous = W > kRedoTol * scan;
    }
    if (sc.hit_refine && !sc.constant_thresh) {
      // Newton steps on F(t) = dist(t) - c0 thr(t), c0 the middle of the
      // band, with a central-difference slope; each step at most one
      // threshold long
      const double c0 = 1.0 - 0.005 * k.quality;
      const double kthr = k.resolution * k.persp / k.quality;
      double t = scan, f = dist - c0 * kthr * t;
      for (int it = 0; it < 8 && fabs(f) > 1e-13 * kthr * t; it++) {
        const double eps = 1e-3 * kthr * t;
        DEInfo bi;
        const double dp = calc_distance_f<double, F, kAlg, IP>(sc.fd, start + dir * (t + eps), detail, sc.N, sc.minN, false, bi);
        const double dm = calc_distance_f<double, F, kAlg, IP>(sc.fd, start + dir * (t - eps), detail, sc.N, sc.minN, false, bi);
        const double g = (dp - dm) / (2.0 * eps) - c0 * kthr;
        if (!(fabs(g) > 0.0)) break;
        double dt = -f / g;
        const double lim = kthr * t;
        dt = dt > lim ? lim : (dt < -lim ? -lim : dt);
        t += dt;
        dist = calc_distance_f<double, F, kAlg, IP>(sc.fd, start + dir * t, detail, sc.N, sc.minN, false, bi);
        cnt.add(bi);
        object = bi.object;
        f = dist - c0 * kthr * t;
      }
      scan = t;
      point = start + dir * scan;
      distThresh = scan * kthr;
    }
  }
  o.point = point;
  o.scan = scan;
  o.distThresh = distThresh;
  o.dist = dist;
  o.buffCount = s.i;
  o.object = object;
  o.found = found;
  o.missed = missed;
}

__device__ __forceinline__ RayState ray_start_state(const DevScene &sc) {
  RayState s;
  s.point = v3<double>(0, 0, 0);
  s.scan = sc.view_distance_min;
  s.step = sc.view_distance_min;
  s.thr = 0.0;
  s.thr0 = 0.0;
  s.drift = 0.0;
  s.dist = 0.0;
  s.i = 0;
  s.object = OBJ_FRACTAL;
  return s;
}

// A ray whose march ended without the FP64 continuation (a miss).
__device__ __forceinline__ void write_ray(const BandBufs &b, int r, const MarchOut &o) {
  b.found[r] = o.found;
  b.object[r] = (uint8_t)o.object;
  b.count[r] = o.buffCount;
  b.depth[r] = o.scan;
  b.thr[r] = o.distThresh;
  b.last[r] = o.dist;
  b.px[r] = o.point.x;
  b.py[r] = o.point.y;
  b.pz[r] = o.point.z;
  if (o.found) {
    const unsigned long long h = atomicAdd(&b.counters[CNT_HITS], 1ull);
    b.hit_list[h] = r;
  }
}

template <class RT>
__device__ __forceinline__ RecPlanes<RT> ray_records(const BandBufs &b, int r, unsigned long long *stats) {
  RecPlanes<RT> rec;
  rec.step = (RT *)b.rec_step;
  rec.dist = (RT *)b.rec_dist;
  rec.thr = (RT *)b.rec_thr;
  rec.ostep = (RT *)b.ovf_step;
  rec.odist = (RT *)b.ovf_dist;
  rec.othr = (RT *)b.ovf_thr;
  rec.it = b.rec_it;
  rec.oit = b.ovf_it;
  rec.ovf_off = b.ovf_off;
  rec.dir = b.ovf_dir;
  rec.dirs = &b.counters[CNT_DIRS];
  rec.chunks = &b.counters[CNT_CHUNKS];
  rec.stats = stats;
  rec.cap = b.cap;
  rec.max_dirs = b.ovf_dirs;
  rec.max_chunks = b.ovf_chunks;
  rec.slab = b.slab;
  rec.r = r;
  return rec;
}

// Cancel polling (mbr_cancel, the reference's programClosed test per pixel,
// src/Render3D.cpp:369-378).  take_work(queue, n) is the hand-out of a band
// work queue (lane 0 only): it returns the first of n items and whether the
// band was stopped.  A hand-out whose range crosses a multiple of
// 2^poll_shift reads the host flag (a read over the bus, so not on every
// hand-out); when set it adds kStopBit to the queue, so every later
// hand-out sees the stop in the value its own atomic returns.
constexpr unsigned long long kStopBit = 1ull << 48;

__device__ __forceinline__ unsigned long long take_work(const BandBufs &b, unsigned long long *queue,
                                                        unsigned long long n, int poll_shift, int &stopped) {
  unsigned long long v = atomicAdd(queue, n);
  if (b.cancel && !(v & kStopBit)) {
    const unsigned long long first = v & (kStopBit - 1);
    if (((first + n) >> poll_shift) != (first >> poll_shift) && *b.cancel != 0) {
      atomicAdd(queue, kStopBit);
      v |= kStopBit;
    }
  }
  stopped = (v & kStopBit) != 0;
  return v & (kStopBit - 1);
}

// Persistent march: one primary ray per lane, up to kMarchIters formula
// iterations per pass, lane states and amortised bookkeeping as in k_ao.
enum : int { MS_NEED = 0, MS_ITER = 1, MS_DONE = 2, MS_EXHAUSTED = 3 };
#ifndef MBR_MARCH_ITERS
#define MBR_MARCH_ITERS 6
#endif
constexpr int kMarchIters = MBR_MARCH_ITERS;

// IP: the integer power of the Mandelbulb when fixed at compile time (8).
template <class T, class RT, int F, bool Plain, int IP>
__device__ __forceinline__ void march_body(const DevScene *__restrict__ scp, BandBufs b, int band_first,
                                           int band_rays, unsigned long long *stats) {
  // integer-power Mandelbulb iterations without inverse trig wherever the
  // records are FP32; the FP64 mode keeps the trigonometric form
  constexpr bool kAlg = sizeof(RT) < 8 || IP != 0;
  const DevScene &sc = *scp;
  // the power-8 variants read a few formula constants in every iteration: a
  // local copy keeps them in registers (the pass's stores and atomics keep
  // the compiler from hoisting loads of the scene block out of the loop)
  const FractalK<T> f_local = (Plain && IP) ? sc.frac<T>() : FractalK<T>{};
  const FractalK<T> &f = (Plain && IP) ? f_local : sc.frac<T>();
  const int N = sc.N, minN = sc.minN;
  const int lane = __lane_id();
  const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const V3<double> start = v3<double>(sc.start[0], sc.start[1], sc.start[2]);

  int state = MS_NEED;
  int r = 0;
  uint32_t pixel = 0;
  V3<double> dir = v3<double>(0, 0, 0);
  RayState s = ray_start_state(sc);
  DECall<T> call;
  T dist = 0;
  int object = OBJ_FRACTAL;
  LaneCount cnt = {0, 0};
  int missed = 0;

  auto begin_de = [&]() {
    march_before_de(sc, start, dir, s);
    DEInfo info;
    state = MS_ITER;
    if (de_call_begin<T, F, Plain, kAlg, IP>(f, cvt<T>(s.point), (T)s.thr, N, minN, false, call, dist, info)) {
      state = MS_DONE;
      object = info.object | (info.iters << 3);
      cnt.de_calls += 1;
    }
  };

  for (;;) {
    const unsigned long long m_need = __ballot(state == MS_NEED);
    const unsigned long long m_iter = __ballot(state == MS_ITER);
    const unsigned long long m_done = __ballot(state == MS_DONE);
    if (!(m_need | m_iter | m_done)) break;
    const int n_need = __popcll(m_need), n_iter = __popcll(m_iter), n_done = __popcll(m_done);

    if (n_need && (4 * n_need >= n_need + n_iter + n_done || n_iter + n_done == 0)) {
      unsigned long long base = 0;
      int stop = 0;
      if (lane == 0) base = take_work(b, &b.counters[CNT_RAYQ], (unsigned long long)n_need, 12, stop);
      base = __shfl(base, 0);
      stop = __shfl(stop, 0);
      if (state == MS_NEED) {
        const unsigned long long mine = base + (unsigned long long)__popcll(m_need & lt_mask);
        if (mine >= (unsigned long long)band_rays) {
          state = MS_EXHAUSTED;
        } else {
          r = (int)mine;
          b.ovf_off[r] = -1;
          const int p = band_first + r;
          const int x = p % sc.width, y = p / sc.width;
          // a cancelled band's remaining rays take the empty record of a
          // cut ray, so later kernels read defined (if meaningless) values
          const bool cut = stop || view_vector(sc, x, y, dir);
          pixel = sc.pixel_base + (uint32_t)p;
          s = ray_start_state(sc);
          if (cut) {
            MarchOut o;
            o.point = v3<double>(0, 0, 0);
            o.scan = 0; o.distThresh = 0; o.dist = 0; o.buffCount = 0; o.object = OBJ_FRACTAL;
            o.found = false; o.missed = false;
            write_ray(b, r, o);
          } else {
            begin_de();
          }
        }
      }
    }

    if (state == MS_ITER) {
      DEInfo info;
      bool fin = false;
#pragma unroll
      for (int k = 0; k < kMarchIters && !fin; k++)
        fin = de_call_step<T, F, Plain, kAlg, IP>(f, call, (T)s.thr, N, minN, false, dist, info);
      if (fin) {
        state = MS_DONE;
        object = info.object | (info.iters << 3);
        cnt.de_calls += 1;
        cnt.iterations += (uint32_t)info.work;
      }
    }

    const unsigned long long m_done2 = __ballot(state == MS_DONE);
    const unsigned long long m_iter2 = __ballot(state == MS_ITER);
    const int n_done2 = __popcll(m_done2), n_iter2 = __popcll(m_iter2);
    if (n_done2 && (4 * n_done2 >= n_done2 + n_iter2 || n_iter2 == 0) && state == MS_DONE) {
      const int out = march_after_de<T, RT>(sc, s, (double)dist, object, pixel, ray_records<RT>(b, r, stats));
      if (out == SO_CONTINUE) {
        begin_de();
      } else if (out == SO_MISS) {
        MarchOut o;
        o.point = s.point;
        o.scan = s.scan;
        o.distThresh = s.thr;
        o.dist = s.dist;
        o.buffCount = s.i;
        o.object = s.object;
        o.found = false;
        o.missed = false;
        write_ray(b, r, o);
        state = MS_NEED;
      } else {
        const unsigned long long h = atomicAdd(&b.counters[CNT_HAND], 1ull);
        b.hand_ray[h] = r;
        b.hand_i[h] = s.i;
        b.hand_scan[h] = s.scan;
        b.hand_step[h] = s.step;
        b.hand_drift[h] = s.drift;
        b.hand_dist[h] = s.dist;
        state = MS_NEED;
      }
    }
  }
  unsigned long long v[ST_COUNT] = {0};
  v[ST_MARCH_DE] = cnt.de_calls;
  v[ST_MARCH_ITERS] = cnt.iterations;
  flush_stats(stats, v);
  (void)missed;
}

template <class T, class RT, int F, bool Plain, int IP = 0>
__global__ void __launch_bounds__(kBlock) k_march(const DevScene *__restrict__ scp, BandBufs b, int band_first,
                                                  int band_rays, unsigned long long *stats) {
  march_body<T, RT, F, Plain, IP>(scp, b, band_first, band_rays, stats);
}


// FP64 continuation of the handed-over rays: one lane per ray, a grid-stride
// loop over the hand-over list whose length the march left in the band
// counters.
template <class RT, int F, int IP = 0>
__global__ void __launch_bounds__(kBlock) k_march_fix(const DevScene *__restrict__ scp, BandBufs b, int band_first,
                                                      unsigned long long *stats) {
  const DevScene &sc = *scp;
  const long long n_hand = (long long)b.counters[CNT_HAND];
  LaneCount cnt = {0, 0};
  unsigned long long hits = 0, missed = 0, redo = 0;
  const V3<double> start = v3<double>(sc.start[0], sc.start[1], sc.start[2]);
  for (long long h = (long long)blockIdx.x * kBlock + threadIdx.x; h < n_hand; h += (long long)gridDim.x * kBlock) {
    const int r = b.hand_ray[h];
    const int p = band_first + r;
    const int x = p % sc.width, y = p / sc.width;
    V3<double> dir;
    view_vector(sc, x, y, dir);
    const uint32_t pixel = sc.pixel_base + (uint32_t)p;
    RayState s = ray_start_state(sc);
    s.i = b.hand_i[h];
    s.scan = b.hand_scan[h];
    s.step = b.hand_step[h];
    s.drift = b.hand_drift[h];
    s.dist = b.hand_dist[h];
    const RecPlanes<RT> rec = ray_records<RT>(b, r, stats);
    MarchOut o;
    bool ambiguous = false;
    march_continue_f64<RT, F, false, IP>(sc, start, dir, pixel, rec, s, o, cnt, ambiguous);
    if (ambiguous) {
      // re-marched from the start by k_march_redo
      b.redo_list[atomicAdd(&b.counters[CNT_REDO], 1ull)] = r;
      redo++;
    } else {
      write_ray(b, r, o);
      hits += o.found;
      missed += o.missed;
    }
  }
  unsigned long long v[ST_COUNT] = {0};
  v[ST_MARCH_DE] = cnt.de_calls;
  v[ST_MARCH_ITERS] = cnt.iterations;
  v[ST_HITS] = hits;
  v[ST_MISSED] = missed;
  v[ST_REDO] = redo;
  flush_stats(stats, v);
}

// FP64 march from the start of the rays k_march_fix listed (FP32_MARCH rays
// whose result FP32 error could have changed).  A kernel of its own: with
// the redo as a second inlined pass of the march loop inside k_march_fix,
// the gfx950 build returned found = false for some lanes whose rays were
// found, depending on which rays shared a wave (the Mandelbox FP32_MARCH
// parity test covers it).
template <class RT, int F>
__global__ void __launch_bounds__(kBlock) k_march_redo(const DevScene *__restrict__ scp, BandBufs b, int band_first,
                                                       unsigned long long *stats) {
  const DevScene &sc = *scp;
  const long long n_redo = (long long)b.counters[CNT_REDO];
  LaneCount cnt = {0, 0};
  unsigned long long hits = 0, missed = 0;
  const V3<double> start = v3<double>(sc.start[0], sc.start[1], sc.start[2]);
  for (long long h = (long long)blockIdx.x * kBlock + threadIdx.x; h < n_redo; h += (long long)gridDim.x * kBlock) {
    const int r = b.redo_list[h];
    const int p = band_first + r;
    const int x = p % sc.width, y = p / sc.width;
    V3<double> dir;
    view_vector(sc, x, y, dir);
    const uint32_t pixel = sc.pixel_base + (uint32_t)p;
    const RecPlanes<RT> rec = ray_records<RT>(b, r, stats);
    MarchOut o;
    bool unused = false;
    march_continue_f64<RT, F>(sc, start, dir, pixel, rec, ray_start_state(sc), o, cnt, unused);
    write_ray(b, r, o);
    hits += o.found;
    missed += o.missed;
  }
  unsigned long long v[ST_COUNT] = {0};
  v[ST_MARCH_DE] = cnt.de_calls;
  v[ST_MARCH_ITERS] = cnt.iterations;
  v[ST_HITS] = hits;
  v[ST_MISSED] = missed;
  flush_stats(stats, v);
}

template <class T>
struct SurfBufs {
  T *base;    // [3][n] mainLight * shade * shadow
  T *spec;    // [3][n] mainLight * specular * shadow
  T *col;     // [3][n] surface colour
  T *fast_ao; // [n]
  T *ao;      // [n * dirs] AO intensity per (hit, direction)
  T *aux;     // [6][n] aux lights: diffuse R, G, B, specular R, G, B
  long long n;
};

template <class T>
__device__ __forceinline__ void load_hit(const DevScene &sc, const BandBufs &b, int r, ShaderIn<T> &in) {
  const ShadeK<T> &k = sc.shade<T>();
  in.point64 = v3<double>(b.px[r], b.py[r], b.pz[r]);
  in.point = cvt<T>(in.point64);
  in.light = v3<T>(k.light[0], k.light[1], k.light[2]);
  in.dist_thresh = (T)b.thr[r];
  in.lastDist = (T)b.last[r];
  in.depth = (T)b.depth[r];
  in.delta = sc.constant_thresh ? (T)(b.depth[r] * sc.sd.resolution * sc.sd.persp) : (T)(b.thr[r] * sc.sd.quality);
  in.detail = in.dist_thresh;
  in.object = b.object[r];
}

// The direction of ray r: the camera's view vector, or the mirrored
// direction of a reflection level.
__device__ __forceinline__ V3<double> ray_dir(const DevScene &sc, const BandBufs &b, int band_first, int r) {
  V3<double> dir;
  if (b.ray_d) {
    dir = v3<double>(b.ray_d[3 * (long long)r + 0], b.ray_d[3 * (long long)r + 1], b.ray_d[3 * (long long)r + 2]);
  } else {
    const int p = band_first + r;
    view_vector(sc, p % sc.width, p / sc.width, dir);
  }
  return dir;
}

template <class T, int F, bool Deep>
__global__ void __launch_bounds__(kBlock) k_surface(const DevScene *__restrict__ scp, BandBufs b, int band_first,
                                                    SurfBufs<T> sb, unsigned long long *stats) {
  const DevScene &sc = *scp;
  const long long n_hits = (long long)b.counters[CNT_HITS];
  LaneCount cnt = {0, 0};
  for (long long tid = (long long)blockIdx.x * kBlock + threadIdx.x; tid < n_hits;
       tid += (long long)gridDim.x * kBlock) {
    const int r = b.hit_list[tid];
    const V3<double> dir = ray_dir(sc, b, band_first, r);
    const ShadeK<T> &k = sc.shade<T>();
    ShaderIn<T> in;
    load_hit(sc, b, r, in);
    in.view = cvt<T>(dir);
    in.normal = normals<T, F, Deep>(sc, in, cnt);
    T sh = dot(in.normal, in.light);
    if (sh < 0) sh = 0;
    const T shade = k.direct_light * ((T(1) - k.shading) + k.shading * sh);
    T shadow = 1;
    if (sc.shadow) shadow = main_shadow<T, F, Deep>(sc, in, cnt);
    V3<T> half = normalize(in.light - in.view);
    T shade2 = dot(in.normal, half);
    if (shade2 < T(0)) shade2 = 0;
    shade2 = Math<T>::pow(shade2, T(30));
    if (shade2 > T(15)) shade2 = 15;
    const T specular = shade2 * k.specular;
    RGBT<T> colour = surface_colour<T, Deep, sizeof(T) == 8 || F == 0>(sc, in, cnt);
    const long long n = sb.n;
    for (int c = 0; c < 3; c++) {
      sb.base[c * n + tid] = k.main_light[c] * shade * shadow;
      sb.spec[c * n + tid] = k.main_light[c] * specular * shadow;
    }
    sb.col[0 * n + tid] = colour.R;
    sb.col[1 * n + tid] = colour.G;
    sb.col[2 * n + tid] = colour.B;
    if (sc.ao && sc.fast_ao) sb.fast_ao[tid] = fast_ao2<T, F, Deep>(sc, in, cnt);
    // AmbientOcclusion estimates with the detail size calcParam holds when
    // it runs: MainShadow's last threshold, float-rounded (src/shaders.cpp:
    // 510-522), when there are shadows; the slow AO pass reads it from here
    if (sc.ao && !sc.fast_ao) sb.fast_ao[tid] = in.detail;
    if constexpr (sizeof(T) == 8) {
      if (sc.env_map) {
        // environment mapping x reflect (ObjectShader, src/shaders.cpp:73-80);
        // k_compose multiplies it by the AO
        double e[3];
        env_mapping(sc, cvt<double>(in.normal), cvt<double>(in.view), e);
        for (int c = 0; c < 3; c++) sb.aux[(6 + c) * n + tid] = (T)(e[c] * sc.reflect);
      }
    }
    if (sc.n_lights > 0 || sc.fake_lights) {
      // k_aux shades the aux (and fake) lights from the normal and the
      // detail size MainShadow left (the aux shadows inherit it)
      sb.aux[0 * n + tid] = in.normal.x;
      sb.aux[1 * n + tid] = in.normal.y;
      sb.aux[2 * n + tid] = in.normal.z;
      sb.aux[3 * n + tid] = in.detail;
    }
    b.col_r[r] = x86_d2u8((double)colour.R * 255);
    b.col_g[r] = x86_d2u8((double)colour.G * 255);
    b.col_b[r] = x86_d2u8((double)colour.B * 255);
  }
  unsigned long long v[ST_COUNT] = {0};
  v[ST_SHADE_DE] = cnt.de_calls;
  v[ST_SHADE_ITERS] = cnt.iterations;
  flush_stats(stats, v);
}

// AuxLightsShader (src/shaders.cpp:1051-1073) per hit, after k_surface: the
// normal and detail size come from k_surface's entries in sb.aux, which are
// replaced by the diffuse and specular sums.  A kernel of its own, so the
// surface kernel of scenes without aux lights keeps its registers.
template <class T, int F, bool Deep>
__global__ void __launch_bounds__(kBlock) k_aux(const DevScene *__restrict__ scp, BandBufs b, int band_first,
                                                SurfBufs<T> sb, unsigned long long *stats) {
  const DevScene &sc = *scp;
  const long long n_hits = (long long)b.counters[CNT_HITS];
  const long long n = sb.n;
  LaneCount cnt = {0, 0};
  for (long long tid = (long long)blockIdx.x * kBlock + threadIdx.x; tid < n_hits;
       tid += (long long)gridDim.x * kBlock) {
    const int r = b.hit_list[tid];
    const V3<double> dir = ray_dir(sc, b, band_first, r);
    ShaderIn<T> in;
    load_hit(sc, b, r, in);
    in.view = cvt<T>(dir);
    in.normal = v3<T>(sb.aux[0 * n + tid], sb.aux[1 * n + tid], sb.aux[2 * n + tid]);
    in.detail = sb.aux[3 * n + tid];
    RGBT<T> al, as;
    aux_lights<T, F, Deep>(sc, in, al, as, cnt);
    if constexpr (sizeof(T) == 8) {
      if (sc.fake_lights) {
        // (auxLights + fakeLights) x colour and auxSpecular + fakeSpecular
        // (ObjectShader, src/shaders.cpp:88-108)
        RGBT<T> fl, fs;
        fake_lights<T>(sc, in, fl, fs, cnt);
        al = RGBT<T>{al.R + fl.R, al.G + fl.G, al.B + fl.B};
        as = RGBT<T>{as.R + fs.R, as.G + fs.G, as.B + fs.B};
      }
    }
    sb.aux[0 * n + tid] = al.R;
    sb.aux[1 * n + tid] = al.G;
    sb.aux[2 * n + tid] = al.B;
    sb.aux[3 * n + tid] = as.R;
    sb.aux[4 * n + tid] = as.G;
    sb.aux[5 * n + tid] = as.B;
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
// Items are (hit, direction) pairs in hit-major order.  A wave walks the
// directions of one hit at a time and takes the next hit index from the
// global counter when they run out, so an item needs no division and a lane
// reloads the hit's data only when its hit changes.
//
// A lane is in one of four states: NEED (wants an item), ITER (advancing a
// distance estimate), DONE (estimate complete, AO step pending) or EXHAUSTED.
// Every pass advances the ITER lanes by up to AOPolicy::iters formula
// iterations (an AO distance estimate of the 1080p power-8 bulb takes 3.6 on
// average).  The AO step logic of DONE lanes runs once 1/done_div of the
// wave waits for it, the item hand-out of NEED lanes once 1/need_div does
// (or nothing else can run), so their cost is shared by many lanes instead
// of being paid on almost every pass.  FP32: 2 iterations, 1/3, 1/8; FP64
// (twice the registers, 2 waves per SIMD): 1 iteration, 1/6, 1/8 (measured
// best of the sweeps on MI355X).
enum : int { AO_NEED = 0, AO_ITER = 1, AO_DONE = 2, AO_EXHAUSTED = 3 };
// Per element type: formula iterations per pass, and the share of the wave
// (1/div) that must wait in DONE or NEED before their block runs.
#ifndef MBR_AO_ITERS
#define MBR_AO_ITERS 2
#endif
#ifndef MBR_AO_DONE_DIV
#define MBR_AO_DONE_DIV 3
#endif
#ifndef MBR_AO_NEED_DIV
#define MBR_AO_NEED_DIV 8
#endif
#ifndef MBR_AO64_ITERS
#define MBR_AO64_ITERS 1
#endif
#ifndef MBR_AO64_DONE_DIV
#define MBR_AO64_DONE_DIV 6
#endif
#ifndef MBR_AO64_NEED_DIV
#define MBR_AO64_NEED_DIV 8
#endif
template <class T>
struct AOPolicy {
  static constexpr bool f64 = sizeof(T) == 8;
  static constexpr int iters = f64 ? MBR_AO64_ITERS : MBR_AO_ITERS;
  static constexpr int done_div = f64 ? MBR_AO64_DONE_DIV : MBR_AO_DONE_DIV;
  static constexpr int need_div = f64 ? MBR_AO64_NEED_DIV : MBR_AO_NEED_DIV;
};

template <class T, int F, bool Deep, bool Plain, int IP = 0>
__device__ __forceinline__ void ao_body(const DevScene *__restrict__ scp, BandBufs b, AODev ao, SurfBufs<T> sb,
                                        unsigned long long *stats) {
  const DevScene &sc = *scp;
  // trig-free integer-power iteration in FP32, and in FP64 for the fixed power
  constexpr bool kAlg = sizeof(T) < 8 || IP != 0;
  const long long n_hits = (long long)b.counters[CNT_HITS];
  unsigned long long *queue = &b.counters[CNT_AOQ];
  // power-8 variants: the formula constants in registers (as in k_march)
  const FractalK<T> f_local = (Plain && IP) ? sc.frac<T>() : FractalK<T>{};
  const FractalK<T> &f = (Plain && IP) ? f_local : sc.frac<T>();
  const ShadeK<T> &k = sc.shade<T>();
  const T *dirs;
  if constexpr (sizeof(T) == 8) dirs = ao.dir_d; else dirs = ao.dir_f;
  const unsigned count = (unsigned)ao.count;
  const int N = sc.N, minN = sc.minN;
  const int lane = __lane_id();
  const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

  // wave-uniform cursor: hit wh, next direction wd
  long long wh = -1;
  unsigned wd = count;
  int state = AO_NEED;
  long long h_cur = -1;  // hit whose data this lane holds
  unsigned long long item = 0;
  V3<T> P = v3<T>(0, 0, 0), dir = P;
  V3<double> P64 = v3<double>(0, 0, 0);
  bool deep = false;  // FP32 path: the hit's detail is below FP32 resolution of its coordinates
  T r = 0, start_dist = 0, end_dist = 1, detail = 0, thr_f = 0;
  T dist = 0;
  int mi = 0, guard = 0;
  DECall<T> call;
  LaneCount cnt = {0, 0};
  unsigned long long ao_steps = 0;

  // start the distance estimate at P + dir * r; DONE when no iteration runs
  auto begin_de = [&]() {
    DEInfo info;
    state = AO_ITER;
    bool done;
    if constexpr (Deep && sizeof(T) < 8) {
      if (deep)
        done = de_call_begin_p64<F>(sc.fd, f, P64 + cvt<double>(dir) * (double)r, detail, N, minN, false, call, dist,
                                    info);
      else
        done = de_call_begin<T, F, Plain, kAlg, IP>(f, P + dir * r, detail, N, minN, false, call, dist,
                                                               info);
    } else {
      done = de_call_begin<T, F, Plain, kAlg, IP>(f, P + dir * r, detail, N, minN, false, call, dist, info);
    }
    if (done) {
      state = AO_DONE;
      mi = info.max_iter;
      cnt.de_calls += 1;
    }
  };

  for (;;) {
    const unsigned long long m_need = __ballot(state == AO_NEED);
    const unsigned long long m_iter = __ballot(state == AO_ITER);
    const unsigned long long m_done = __ballot(state == AO_DONE);
    if (!(m_need | m_iter | m_done)) break;
    const int n_need = __popcll(m_need), n_iter = __popcll(m_iter), n_done = __popcll(m_done);

    if (n_need && (AOPolicy<T>::need_div * n_need >= n_need + n_iter + n_done || n_iter + n_done == 0)) {
      // hand out the next items to the lanes that need one, in lane order
      const unsigned n = (unsigned)n_need;
      const unsigned off = (unsigned)__popcll(m_need & lt_mask);
      const unsigned avail = count - wd;
      long long h = wh;
      unsigned d = wd + off;
      if (n > avail) {
        // new hits for the lanes beyond this hit's remaining directions
        const unsigned extra = n - avail;
        const unsigned khits = (extra + count - 1) / count;
        unsigned long long base = 0;
        if (lane == 0) {
          int stop = 0;
          base = take_work(b, queue, (unsigned long long)khits, 8, stop);
          // cancelled: no further hits (their AO intensities stay unwritten)
          if (stop) base = (unsigned long long)n_hits;
        }
        base = __shfl(base, 0);
        if (off >= avail) {
          const unsigned j = off - avail;
          const unsigned q = count >= 64 ? 0u : j / count;  // several hits only when count < 64
          h = (long long)base + q;
          d = j - q * count;
        }
        wh = (long long)base + (khits - 1);
        wd = extra - (khits - 1) * count;
      } else {
        wd += n;
      }
      if (state == AO_NEED) {
        if (h >= n_hits) {
          state = AO_EXHAUSTED;
        } else {
          if (h != h_cur) {
            h_cur = h;
            ShaderIn<T> in;
            load_hit(sc, b, b.hit_list[h], in);
            P = in.point;
            P64 = in.point64;
            if constexpr (Deep) deep = (double)in.detail < 2.0 * kMixedSwitch * length(P64);
            start_dist = in.delta;
            end_dist = in.delta / k.resolution;
            detail = sb.fast_ao[h];  // the detail size k_surface left (after MainShadow)
            thr_f = (T)(float)in.dist_thresh;  // float dist_thresh (src/shaders.cpp:642)
          }
          item = (unsigned long long)h * count + d;
          dir = v3<T>(dirs[4 * d + 0], dirs[4 * d + 1], dirs[4 * d + 2]);
          r = start_dist;
          guard = 0;
          if (r < end_dist) begin_de();
          else sb.ao[item] = T(1);
        }
      }
    }

    if (state == AO_ITER) {
      DEInfo info;
      bool fin = false;
#pragma unroll
      for (int k = 0; k < AOPolicy<T>::iters && !fin; k++)
        fin = de_call_step<T, F, Plain, kAlg, IP>(f, call, detail, N, minN, false, dist, info);
      if (fin) {
        state = AO_DONE;
        mi = info.max_iter;
        cnt.de_calls += 1;
        cnt.iterations += (uint32_t)info.work;
      }
    }

    const unsigned long long m_done2 = __ballot(state == AO_DONE);
    const unsigned long long m_iter2 = __ballot(state == AO_ITER);
    const int n_done2 = __popcll(m_done2), n_iter2 = __popcll(m_iter2);
    if (n_done2 && (AOPolicy<T>::done_div * n_done2 >= n_done2 + n_iter2 || n_iter2 == 0) && state == AO_DONE) {
      // one step of the AO ray (src/shaders.cpp:625-662)
      ao_steps++;
      if (dist < thr_f || mi) {
        T shadowTemp = T(1) - (end_dist - r) / end_dist;
        if (shadowTemp < T(0)) shadowTemp = 0;
        sb.ao[item] = shadowTemp;
        state = AO_NEED;
      } else {
        r += dist * T(2);
        guard++;
        if (!(r < end_dist) || guard >= kMaxShadeSteps) {
          sb.ao[item] = T(1);
          state = AO_NEED;
        } else {
          begin_de();
        }
      }
    }
  }
  unsigned long long v[ST_COUNT] = {0};
  v[ST_SHADE_DE] = cnt.de_calls;
  v[ST_SHADE_ITERS] = cnt.iterations;
  v[ST_AO_DE] = cnt.de_calls;
  v[ST_AO_ITERS] = cnt.iterations;
  v[ST_AO_STEPS] = ao_steps;
  flush_stats(stats, v);
}

template <class T, int F, bool Deep, bool Plain, int IP = 0>
__global__ void __launch_bounds__(kBlock) k_ao(const DevScene *__restrict__ scp, BandBufs b, AODev ao,
                                               SurfBufs<T> sb, unsigned long long *stats) {
  ao_body<T, F, Deep, Plain, IP>(scp, b, ao, sb, stats);
}

// The FP32 plain variant (the benchmarked AO pass) held to 6 waves per SIMD.
template <class T, int F, int IP = 0>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6, 8)))
    k_ao_f32_plain(const DevScene *__restrict__ scp, BandBufs b, AODev ao, SurfBufs<T> sb,
                   unsigned long long *stats) {
  ao_body<T, F, false, true, IP>(scp, b, ao, sb, stats);
}

template <class T, class RT>
__global__ void __launch_bounds__(kBlock) k_compose(const DevScene *__restrict__ scp, BandBufs b, AODev ao,
                                                    SurfBufs<T> sb) {
  const DevScene &sc = *scp;
  const long long n_hits = (long long)b.counters[CNT_HITS];
  const ShadeK<T> &k = sc.shade<T>();
  const long long n = sb.n;
  const int lane = __lane_id();
  const long long wave0 = ((long long)blockIdx.x * kBlock + threadIdx.x) / 64;
  const long long waves = (long long)gridDim.x * (kBlock / 64);
  // one wave per hit: the lanes read the hit's AO intensities coalesced,
  // lane j summing directions j, j + 64, ... in order, then the 64 partial
  // sums are added pairwise (a fixed order, so results are reproducible)
  for (long long tid = wave0; tid < n_hits; tid += waves) {
    const int r = b.hit_list[tid];
    RGBT<T> amb = {0, 0, 0};
    if (sc.ao) {
      if (sc.fast_ao) {
        const T a = sb.fast_ao[tid];
        amb = RGBT<T>{a, a, a};
      } else {
        const T *it = sb.ao + tid * ao.count;
        for (int v = lane; v < ao.count; v += 64) {
          const T s = it[v];
          amb.R += s * T(ao.rgb[4 * v + 0]);
          amb.G += s * T(ao.rgb[4 * v + 1]);
          amb.B += s * T(ao.rgb[4 * v + 2]);
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
          amb.R += __shfl_xor(amb.R, off);
          amb.G += __shfl_xor(amb.G, off);
          amb.B += __shfl_xor(amb.B, off);
        }
        const T den = T(ao.count) * T(256);
        amb.R /= den;
        amb.G /= den;
        amb.B /= den;
      }
    }
    // environment mapping x reflect, times the AO (0 without it)
    RGBT<T> env = {0, 0, 0};
    if (sc.env_map) env = RGBT<T>{sb.aux[6 * n + tid], sb.aux[7 * n + tid], sb.aux[8 * n + tid]};
    if (lane == 0 && b.sp_r) {
      // reflections: the object part and the specular part apart
      const T cR = sb.col[0 * n + tid], cG = sb.col[1 * n + tid], cB = sb.col[2 * n + tid];
      T oR = env.R * amb.R + (amb.R * k.global_ilum + sb.base[0 * n + tid]) * cR;
      T oG = env.G * amb.G + (amb.G * k.global_ilum + sb.base[1 * n + tid]) * cG;
      T oB = env.B * amb.B + (amb.B * k.global_ilum + sb.base[2 * n + tid]) * cB;
      T sR = sb.spec[0 * n + tid], sG = sb.spec[1 * n + tid], sB = sb.spec[2 * n + tid];
      if (sc.n_lights > 0 || sc.fake_lights) {
        oR += sb.aux[0 * n + tid] * cR;
        oG += sb.aux[1 * n + tid] * cG;
        oB += sb.aux[2 * n + tid] * cB;
        sR = sb.aux[3 * n + tid] + sR;
        sG = sb.aux[4 * n + tid] + sG;
        sB = sb.aux[5 * n + tid] + sB;
      }
      ((RT *)b.sh_r)[r] = (RT)oR;
      ((RT *)b.sh_g)[r] = (RT)oG;
      ((RT *)b.sh_b)[r] = (RT)oB;
      ((RT *)b.sp_r)[r] = (RT)sR;
      ((RT *)b.sp_g)[r] = (RT)sG;
      ((RT *)b.sp_b)[r] = (RT)sB;
    } else if (lane == 0) {
      const T cR = sb.col[0 * n + tid], cG = sb.col[1 * n + tid], cB = sb.col[2 * n + tid];
      if (sc.n_lights > 0 || sc.fake_lights) {
        // (ambient + main light) x colour + aux x colour, then aux specular
        // + main specular (ObjectShader, src/shaders.cpp:96-108)
        ((RT *)b.sh_r)[r] = (RT)(((env.R * amb.R + (amb.R * k.global_ilum + sb.base[0 * n + tid]) * cR) +
                                  sb.aux[0 * n + tid] * cR) +
                                 (sb.aux[3 * n + tid] + sb.spec[0 * n + tid]));
        ((RT *)b.sh_g)[r] = (RT)(((env.G * amb.G + (amb.G * k.global_ilum + sb.base[1 * n + tid]) * cG) +
                                  sb.aux[1 * n + tid] * cG) +
                                 (sb.aux[4 * n + tid] + sb.spec[1 * n + tid]));
        ((RT *)b.sh_b)[r] = (RT)(((env.B * amb.B + (amb.B * k.global_ilum + sb.base[2 * n + tid]) * cB) +
                                  sb.aux[2 * n + tid] * cB) +
                                 (sb.aux[5 * n + tid] + sb.spec[2 * n + tid]));
      } else {
        ((RT *)b.sh_r)[r] =
            (RT)((env.R * amb.R + (amb.R * k.global_ilum + sb.base[0 * n + tid]) * cR) + sb.spec[0 * n + tid]);
        ((RT *)b.sh_g)[r] =
            (RT)((env.G * amb.G + (amb.G * k.global_ilum + sb.base[1 * n + tid]) * cG) + sb.spec[1 * n + tid]);
        ((RT *)b.sh_b)[r] =
            (RT)((env.B * amb.B + (amb.B * k.global_ilum + sb.base[2 * n + tid]) * cB) + sb.spec[2 * n + tid]);
      }
    }
  }
}

// The stages of VolumetricShader that run distance estimates, at a step
// group's point: the volumetric lights (src/shaders.cpp:314-341; stage 0)
// and the iteration fog (:385-465; stage 1), whose record's iteration count
// sets the opacity and the fog colour and whose group is lit by the main
// light and the enabled aux slots through their shadows, plus AO.  `pin` is
// the shader input of the ray's hit, moved to each group's point and
// threshold; b and r name the ray's step records.
template <class T, int F>
struct FxStages {
  const DevScene &sc;
  const BandBufs &b;
  int r;
  ShaderIn<T> &pin;
  LaneCount &cnt;
  const AODev &ao;

  __device__ void iter_fog(int index, T step, VolOut<T> &o, float &totalOpacity) const {
    int L = 0;
    if (index < b.slab) {
      L = b.rec_it[(long long)index * b.cap + r];
    } else {
      const long long dj = b.ovf_off[r];
      const int ch = dj >= 0 ? b.ovf_dir[dj * kDirSlots + (index - b.slab) / kChunk] : -1;
      if (ch >= 0) L = b.ovf_it[(long long)ch * kChunk + (index - b.slab) % kChunk];
    }
    T opacity = iter_opacity<T>(step, T(L), T(sc.N), (T)sc.iter_fog_trim, (T)sc.iter_fog_opacity);
    if (!(opacity > T(0))) return;
    const T trim = (T)sc.iter_fog_trim;
    const T iterFactor = T(2) * (T(L) - trim) / (T(sc.N) - trim);
    T kk = iterFactor;
    if (kk > T(1)) kk = 1;
    if (kk < T(0)) kk = 0;
    T kn = T(1) - kk;
    T fR = T(sc.fog1[0]) * kn + T(sc.fog2[0]) * kk;
    T fG = T(sc.fog1[1]) * kn + T(sc.fog2[1]) * kk;
    T fB = T(sc.fog1[2]) * kn + T(sc.fog2[2]) * kk;
    T k2 = iterFactor - T(1);
    if (k2 < T(0)) k2 = 0;
    if (k2 > T(1)) k2 = 1;
    kn = T(1) - k2;
    fR = fR * kn + T(sc.fog3[0]) * k2;
    fG = fG * kn + T(sc.fog3[1]) * k2;
    fB = fB * kn + T(sc.fog3[2]) * k2;
    const ShadeK<T> &k = sc.shade<T>();
    RGBT<T> nc = {0, 0, 0};
    const T mi = (T)sc.main_light_intensity, dl = k.direct_light;
    if (mi * dl > T(0)) {
      const T sh = main_shadow<T, F, false>(sc, pin, cnt);
      nc.R += sh * T(sc.main_light_colour[0]) / T(65536) * mi * dl;
      nc.G += sh * T(sc.main_light_colour[1]) / T(65536) * mi * dl;
      nc.B += sh * T(sc.main_light_colour[2]) / T(65536) * mi * dl;
    }
    for (int s4 = 0; s4 < 4; s4++) {
      if (!sc.aux_slot_on[s4]) continue;
      const double *Lt = sc.vol_aux[s4];
      const V3<T> d = v3<T>((T)Lt[0], (T)Lt[1], (T)Lt[2]) - pin.point;
      const T dist = length(d);
      const T d2 = dist * dist;
      const V3<T> lv = normalize(d);
      const T light = aux_shadow<T, F, false>(sc, pin, dist, lv, cnt);
      const T intensity = (T)Lt[3] * T(100);
      nc.R += light * T(Lt[4]) / T(65536) / d2 * intensity;
      nc.G += light * T(Lt[5]) / T(65536) / d2 * intensity;
      nc.B += light * T(Lt[6]) / T(65536) / d2 * intensity;
    }
    if (sc.ao && !sc.fast_ao) {
      AOTable<T> tab;
      if constexpr (sizeof(T) == 8) tab.dir = ao.dir_d; else tab.dir = ao.dir_f;
      tab.rgb = ao.rgb;
      tab.count = ao.count;
      const RGBT<T> a = ambient_occlusion<T, F, false>(sc, pin, tab, cnt);
      nc.R += a.R * k.global_ilum;
      nc.G += a.G * k.global_ilum;
      nc.B += a.B * k.global_ilum;
    }
    if (opacity > T(1)) opacity = 1;
    o.rgb.R = o.rgb.R * (T(1) - opacity) + nc.R * opacity * fR / T(65536);
    o.rgb.G = o.rgb.G * (T(1) - opacity) + nc.G * opacity * fG / T(65536);
    o.rgb.B = o.rgb.B * (T(1) - opacity) + nc.B * opacity * fB / T(65536);
    totalOpacity = (float)(opacity + (T(1) - opacity) * T(totalOpacity));
    o.alpha = opacity + (T(1) - opacity) * o.alpha;
  }

  __device__ void operator()(int stage, int index, V3<double> p64, T thr, T step, VolOut<T> &o,
                             float &totalOpacity) const {
    pin.point64 = p64;
    pin.point = cvt<T>(p64);
    pin.dist_thresh = thr;
    if (stage == 1) {
      if (sc.iter_fog) iter_fog(index, step, o, totalOpacity);
      return;
    }
    if constexpr (sizeof(T) == 8) {
      if (sc.fake_lights) {
#pragma clang fp contract(off)
        // fake lights' glow (src/shaders.cpp:302-312); 1e-30f is the
        // reference's float literal
        int L = 0;
        T rt = compute<T, MODE_ORBIT_TRAP>(sc.frac<T>(), pin.point, sc.N, &L);
        cnt.de_calls += 1;
        rt = Math<T>::sqrt(T(1) / (rt + T(1.0e-30f)));
        const T e = T(10) / T(sc.fl_visibility_size);
        const T glow = T(1) / (Math<T>::pow(rt, e) * Math<T>::pow(T(10), e) + T(1e-100));
        const T v = T(sc.fl_visibility);
        o.rgb.R += glow * step * v;
        o.rgb.G += glow * step * v;
        o.rgb.B += glow * step * v;
        o.alpha += glow * step * v;
      }
    }
    if (sc.vol_main) {
      const T sh = main_shadow<T, F, false>(sc, pin, cnt);
      const T I0 = (T)sc.vol_intensity[0];
      o.rgb.R += sh * step * I0 * T(sc.main_light_colour[0]) / T(65536);
      o.rgb.G += sh * step * I0 * T(sc.main_light_colour[1]) / T(65536);
      o.rgb.B += sh * step * I0 * T(sc.main_light_colour[2]) / T(65536);
      o.alpha += (sh + sh + sh) / T(3) * step * I0;
    }
    for (int k = 0; k < 4; k++) {
      if (!sc.vol_aux_on[k]) continue;
      const double *L = sc.vol_aux[k];
      const V3<T> d = v3<T>((T)L[0], (T)L[1], (T)L[2]) - pin.point;
      const T dl = length(d);
      const T d2 = dl * dl;
      const V3<T> lv = normalize(d);
      const T light = aux_shadow<T, F, false>(sc, pin, dl, lv, cnt);
      const T Ik = (T)sc.vol_intensity[k + 1];
      o.rgb.R += light * T(L[4]) / T(65536) * Ik * step / d2;
      o.rgb.G += light * T(L[5]) / T(65536) * Ik * step / d2;
      o.rgb.B += light * T(L[6]) / T(65536) * Ik * step / d2;
      o.alpha += light * Ik * step / d2;
    }
  }
};

// ---------------------------------------------------------------- finish
// One pixel: background or shaded surface, then VolumetricShader over the
// ray's step records.  FX: the volumetric lights' shadow marches run here
// (distance estimates of formula F).
template <class T, class RT, int F, bool FX, bool VIS>
__device__ __forceinline__ void finish_body(const DevScene &sc, const BandBufs &b, int band_first, int r,
                                            const FrameOut &out, LaneCount &cnt, const AODev &ao) {
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
  const RT *st = (const RT *)b.rec_step, *di = (const RT *)b.rec_dist, *th = (const RT *)b.rec_thr;
  const long long cap = b.cap;
  const int slab = b.slab;
  // visible aux lights glow along the ray: its direction and the scan of
  // the last record (the sum of all recorded steps)
  VisRay vray;
  const VisRay *vr = nullptr;
  const bool vis = VIS && sc.n_lights > 0 && sc.sd.aux_visibility > 0;
  if (vis || FX) {
    const int x = p % sc.width, y = p / sc.width;
    view_vector(sc, x, y, vray.dir);
    vray.start = v3<double>(sc.start[0], sc.start[1], sc.start[2]);
    vr = &vray;
  }
  // the volumetric lights of a step group (src/shaders.cpp:314-341), from
  // the shader input of the pixel with the group's point and threshold
  ShaderIn<T> pin;
  if constexpr (FX) {
    load_hit(sc, b, r, pin);
    pin.view = cvt<T>(vray.dir);
  }
  auto fx = [&](int stage, int index, V3<double> p64, T thr, T step, VolOut<T> &o, float &totalOpacity) {
    if constexpr (FX) {
      FxStages<T, F>{sc, b, r, pin, cnt, ao}(stage, index, p64, thr, step, o, totalOpacity);
    } else {
      (void)stage; (void)index; (void)p64; (void)thr; (void)step; (void)o; (void)totalOpacity;
    }
  };
  if (n <= slab) {
    auto rec = [&](int i, T &s, T &d, T &t) {
      const long long o = (long long)i * cap + r;
      s = (T)st[o]; d = (T)di[o]; t = (T)th[o];
    };
    if (vr) {
      double sum = 0.0;
      for (int i = 0; i < n; i++) sum += (double)st[(long long)i * cap + r];
      vray.scan_end = sum;
    }
    v = volumetric<T>(sc, depth, n, px, alpha, rec, vr, vis, fx);
  } else {
    // records from the slab on sit in the chunks the ray's directory lists
    // (the host renders a band again when a record found the pool full)
    const RT *ost = (const RT *)b.ovf_step, *odi = (const RT *)b.ovf_dist, *oth = (const RT *)b.ovf_thr;
    const long long dj = b.ovf_off[r];
    const int *dir = b.ovf_dir + (dj >= 0 ? dj : 0) * kDirSlots;
    auto rec = [&](int i, T &s, T &d, T &t) {
      const int ch = (i >= slab && dj >= 0) ? dir[(i - slab) / kChunk] : -1;
      if (i < slab) {
        const long long o = (long long)i * cap + r;
        s = (T)st[o]; d = (T)di[o]; t = (T)th[o];
      } else if (ch >= 0) {
        const long long o = (long long)ch * kChunk + (i - slab) % kChunk;
        s = (T)ost[o]; d = (T)odi[o]; t = (T)oth[o];
      } else {
        s = 0; d = 0; t = 0;
      }
    };
    if (vr) {
      double sum = 0.0;
      for (int i = 0; i < n; i++) {
        T s, d, t;
        rec(i, s, d, t);
        sum += (double)s;
      }
      vray.scan_end = sum;
    }
    v = volumetric<T>(sc, depth, n, px, alpha, rec, vr, vis, fx);
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

// VIS: the variant with the glow of visible aux lights
template <class T, class RT, bool VIS>
__global__ void __launch_bounds__(kBlock) k_finish(const DevScene *__restrict__ scp, BandBufs b, int band_first,
                                                   int band_rays, FrameOut out) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= band_rays) return;
  LaneCount cnt = {0, 0};
  finish_body<T, RT, 0, false, VIS>(*scp, b, band_first, r, out, cnt, AODev{});
}

// With volumetric lights: a grid-stride loop (the shadow marches make the
// per-pixel cost vary widely), distance-estimate counts into the stats.
template <class T, class RT, int F>
__global__ void __launch_bounds__(kBlock) k_finish_fx(const DevScene *__restrict__ scp, BandBufs b, int band_first,
                                                      int band_rays, FrameOut out, unsigned long long *stats,
                                                      AODev ao) {
  LaneCount cnt = {0, 0};
  for (int r = blockIdx.x * kBlock + threadIdx.x; r < band_rays; r += gridDim.x * kBlock)
    finish_body<T, RT, F, true, true>(*scp, b, band_first, r, out, cnt, ao);
  unsigned long long v[ST_COUNT] = {0};
  v[ST_SHADE_DE] = cnt.de_calls;
  v[ST_SHADE_ITERS] = cnt.iterations;
  flush_stats(stats, v);
}

// ----------------------------------------------------------- reflections
__global__ void k_refl_clear(uint8_t *active, int n) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r < n) active[r] = 0;
}

// The mirrored rays of the reflecting hits of `prev` (src/Render3D.cpp:
// 478-498): normal at the hit (CalculateNormals with the hit's threshold as
// detail size), view vector mirrored about it, start one threshold along it.
template <int F>
__global__ void __launch_bounds__(kBlock) k_refl_setup(const DevScene *__restrict__ scp, BandBufs prev, BandBufs next,
                                                       int band_first, unsigned long long *stats) {
  const DevScene &sc = *scp;
  const long long n_hits = (long long)prev.counters[CNT_HITS];
  LaneCount cnt = {0, 0};
  for (long long t = (long long)blockIdx.x * kBlock + threadIdx.x; t < n_hits; t += (long long)gridDim.x * kBlock) {
    const int r = prev.hit_list[t];
    if (sc.refl[prev.object[r] % 7] == 0.0) continue;
    const V3<double> dir = ray_dir(sc, prev, band_first, r);
    ShaderIn<double> in;
    load_hit(sc, prev, r, in);
    in.view = dir;
    const V3<double> n = normals<double, F, false>(sc, in, cnt);
    const V3<double> d2 = dir - (n * dot(dir, n)) * 2.0;
    const double thr = prev.thr[r];
    const V3<double> o = in.point64 + d2 * thr;
    next.ray_o[3 * (long long)r + 0] = o.x;
    next.ray_o[3 * (long long)r + 1] = o.y;
    next.ray_o[3 * (long long)r + 2] = o.z;
    next.ray_d[3 * (long long)r + 0] = d2.x;
    next.ray_d[3 * (long long)r + 1] = d2.y;
    next.ray_d[3 * (long long)r + 2] = d2.z;
    next.ray_t0[r] = thr;
    next.active[r] = 1;
    next.hand_ray[atomicAdd(&next.counters[CNT_HAND], 1ull)] = r;
  }
  unsigned long long v[ST_COUNT] = {0};
  v[ST_SHADE_DE] = cnt.de_calls;
  v[ST_SHADE_ITERS] = cnt.iterations;
  flush_stats(stats, v);
}

// The FP64 march (RayMarching with its binary search) of a reflection
// level's rays, with the same step jitter as the pixel's primary ray.
template <int F>
__global__ void __launch_bounds__(kBlock) k_refl_march(const DevScene *__restrict__ scp, BandBufs b, int band_first,
                                                       unsigned long long *stats) {
  const DevScene &sc = *scp;
  const long long n_rays = (long long)b.counters[CNT_HAND];
  LaneCount cnt = {0, 0};
  unsigned long long hits = 0, missed = 0;
  for (long long h = (long long)blockIdx.x * kBlock + threadIdx.x; h < n_rays; h += (long long)gridDim.x * kBlock) {
    const int r = b.hand_ray[h];
    b.ovf_off[r] = -1;
    const V3<double> start = v3<double>(b.ray_o[3 * (long long)r], b.ray_o[3 * (long long)r + 1],
                                        b.ray_o[3 * (long long)r + 2]);
    const V3<double> dir = v3<double>(b.ray_d[3 * (long long)r], b.ray_d[3 * (long long)r + 1],
                                      b.ray_d[3 * (long long)r + 2]);
    const uint32_t pixel = sc.pixel_base + (uint32_t)(band_first + r);
    RayState s = ray_start_state(sc);
    s.thr0 = b.ray_t0[r];
    MarchOut o;
    bool ambiguous = false;
    march_continue_f64<double, F, true>(sc, start, dir, pixel, ray_records<double>(b, r, stats), s, o, cnt,
                                        ambiguous);
    write_ray(b, r, o);
    hits += o.found;
    missed += o.missed;
  }
  unsigned long long v[ST_COUNT] = {0};
  v[ST_MARCH_DE] = cnt.de_calls;
  v[ST_MARCH_ITERS] = cnt.iterations;
  v[ST_HITS] = hits;
  v[ST_MISSED] = missed;
  flush_stats(stats, v);
}

// Per pixel: the levels from the last one reached down to the primary ray
// (src/Render3D.cpp:501-571): each shaded surface (or background) blended
// with the next level's result by its reflection coefficient, then that
// level's VolumetricShader over its own step records.  FX: with volumetric
// lights or iteration fog, each level's VolumetricShader runs their shadow
// marches (formula F) from that level's hit, as the primary ray's does.
template <int F, bool VIS, bool FX>
__global__ void __launch_bounds__(kBlock) k_refl_finish(const DevScene *__restrict__ scp,
                                                        const BandBufs *__restrict__ levels, int n_levels,
                                                        int band_first, int band_rays, FrameOut out, AODev ao,
                                                        unsigned long long *stats) {
  const DevScene &sc = *scp;
  LaneCount cnt = {0, 0};
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r < band_rays) {
  const int p = band_first + r;
  int rayEnd = 0;
  for (int k = 1; k < n_levels; k++) {
    if (!levels[k].active[r]) break;
    rayEnd = k;
  }
  const int maxRay = n_levels - 1;
  RGBT<double> result = {0, 0, 0};
  double opacity0 = 0, alpha0 = 0;
  for (int ray = rayEnd; ray >= 0; ray--) {
    const BandBufs &b = levels[ray];
    const bool found = b.found[r];
    const V3<double> dir = ray_dir(sc, b, band_first, r);
    RGBT<double> obj = {0, 0, 0}, bg = {0, 0, 0}, spec = {0, 0, 0};
    if (found) {
      obj = RGBT<double>{((const double *)b.sh_r)[r], ((const double *)b.sh_g)[r], ((const double *)b.sh_b)[r]};
      spec = RGBT<double>{((const double *)b.sp_r)[r], ((const double *)b.sp_g)[r], ((const double *)b.sp_b)[r]};
    } else {
      bg = background<double>(sc, dir);
    }
    const double reflect = sc.refl[b.object[r] % 7];
    RGBT<double> px;
    if (maxRay > 0 && rayEnd > 0 && ray != rayEnd) {
      px.R = result.R * reflect + (1.0 - reflect) * (obj.R + bg.R) + spec.R;
      px.G = result.G * reflect + (1.0 - reflect) * (obj.G + bg.G) + spec.G;
      px.B = result.B * reflect + (1.0 - reflect) * (obj.B + bg.B) + spec.B;
    } else {
      px.R = obj.R + bg.R + spec.R;
      px.G = obj.G + bg.G + spec.G;
      px.B = obj.B + bg.B + spec.B;
    }
    const double alpha = found ? 1.0 : 0.0;
    const int n = b.count[r];
    const double depth = b.depth[r];
    const double *st = (const double *)b.rec_step, *di = (const double *)b.rec_dist, *th = (const double *)b.rec_thr;
    const double *ost = (const double *)b.ovf_step, *odi = (const double *)b.ovf_dist, *oth = (const double *)b.ovf_thr;
    const long long cap = b.cap;
    const int slab = b.slab;
    const long long dj = b.ovf_off[r];
    auto rec = [&](int i, double &s, double &d, double &t) {
      if (i < slab) {
        const long long o = (long long)i * cap + r;
        s = st[o]; d = di[o]; t = th[o];
        return;
      }
      const int ch = dj >= 0 ? b.ovf_dir[dj * kDirSlots + (i - slab) / kChunk] : -1;
      if (ch >= 0) {
        const long long o = (long long)ch * kChunk + (i - slab) % kChunk;
        s = ost[o]; d = odi[o]; t = oth[o];
      } else {
        s = 0; d = 0; t = 0;
      }
    };
    VisRay vray;
    const VisRay *vr = nullptr;
    const bool vis = VIS && sc.n_lights > 0 && sc.sd.aux_visibility > 0;
    if (vis || FX) {
      vray.dir = dir;
      vray.start = ray == 0 ? v3<double>(sc.start[0], sc.start[1], sc.start[2])
                            : v3<double>(b.ray_o[3 * (long long)r], b.ray_o[3 * (long long)r + 1],
                                         b.ray_o[3 * (long long)r + 2]);
      double sum = 0.0;
      for (int i = 0; i < n; i++) {
        double s_, d_, t_;
        rec(i, s_, d_, t_);
        sum += s_;
      }
      vray.scan_end = sum;
      vr = &vray;
    }
    ShaderIn<double> pin;
    if constexpr (FX) {
      load_hit(sc, b, r, pin);
      pin.view = dir;
    }
    auto fx = [&](int stage, int index, V3<double> p64, double thr, double step, VolOut<double> &o,
                  float &totalOpacity) {
      if constexpr (FX) {
        FxStages<double, F>{sc, b, r, pin, cnt, ao}(stage, index, p64, thr, step, o, totalOpacity);
      } else {
        (void)stage; (void)index; (void)p64; (void)thr; (void)step; (void)o; (void)totalOpacity;
      }
    };
    const VolOut<double> v = volumetric<double>(sc, depth, n, px, alpha, rec, vr, vis, fx);
    result = v.rgb;
    if (ray == 0) {
      opacity0 = v.opacity;
      alpha0 = v.alpha;
    }
  }
  const BandBufs &b0 = levels[0];
  const bool found0 = b0.found[r];
  if (out.rgb[0]) out.rgb[0][p] = (float)result.R;
  if (out.rgb[1]) out.rgb[1][p] = (float)result.G;
  if (out.rgb[2]) out.rgb[2][p] = (float)result.B;
  if (out.z) out.z[p] = found0 ? (float)b0.depth[r] : 1e20f;
  if (out.colour[0]) out.colour[0][p] = found0 ? b0.col_r[r] : 0;
  if (out.colour[1]) out.colour[1][p] = found0 ? b0.col_g[r] : 0;
  if (out.colour[2]) out.colour[2][p] = found0 ? b0.col_b[r] : 0;
  if (out.opacity) out.opacity[p] = x86_d2u16(opacity0 * 65535.0);
  if (out.alpha) out.alpha[p] = x86_d2u16(alpha0 * 65535.0);
  }
  if constexpr (FX) {
    unsigned long long v[ST_COUNT] = {0};
    v[ST_SHADE_DE] = cnt.de_calls;
    v[ST_SHADE_ITERS] = cnt.iterations;
    flush_stats(stats, v);
  }
}

// ---------------------------------------------------------------- launch
static inline dim3 grid_for(int n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

hipError_t launch_refl_clear(const BandBufs &next, int band_rays, hipStream_t s) {
  if (band_rays <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_refl_clear, grid_for(band_rays), dim3(kBlock), 0, s, next.active, band_rays);
  return hipGetLastError();
}

hipError_t launch_refl_setup(const DevScene *sc, const BandBufs &prev, const BandBufs &next, int band_first,
                             int formula, int blocks, unsigned long long *stats, hipStream_t s) {
  const dim3 grid((unsigned)std::max(1, blocks));
  if (formula_generic(formula))
    hipLaunchKernelGGL((k_refl_setup<0>), grid, dim3(kBlock), 0, s, sc, prev, next, band_first, stats);
  else if (formula == 7)
    hipLaunchKernelGGL((k_refl_setup<7>), grid, dim3(kBlock), 0, s, sc, prev, next, band_first, stats);
  else if (formula == 8)
    hipLaunchKernelGGL((k_refl_setup<8>), grid, dim3(kBlock), 0, s, sc, prev, next, band_first, stats);
  else
    hipLaunchKernelGGL((k_refl_setup<2>), grid, dim3(kBlock), 0, s, sc, prev, next, band_first, stats);
  return hipGetLastError();
}

hipError_t launch_refl_march(const DevScene *sc, const BandBufs &next, int band_first, int formula, int blocks,
                             unsigned long long *stats, hipStream_t s) {
  const dim3 grid((unsigned)std::max(1, blocks));
  if (formula_generic(formula))
    hipLaunchKernelGGL((k_refl_march<0>), grid, dim3(kBlock), 0, s, sc, next, band_first, stats);
  else if (formula == 7)
    hipLaunchKernelGGL((k_refl_march<7>), grid, dim3(kBlock), 0, s, sc, next, band_first, stats);
  else if (formula == 8)
    hipLaunchKernelGGL((k_refl_march<8>), grid, dim3(kBlock), 0, s, sc, next, band_first, stats);
  else
    hipLaunchKernelGGL((k_refl_march<2>), grid, dim3(kBlock), 0, s, sc, next, band_first, stats);
  return hipGetLastError();
}

hipError_t launch_refl_finish(const DevScene *sc, const BandBufs *levels, int n_levels, int band_first,
                              int band_rays, FrameOut out, int vis, int fx, int formula, AODev ao,
                              unsigned long long *stats, hipStream_t s) {
  if (band_rays <= 0) return hipSuccess;
  const dim3 grid = grid_for(band_rays);
  auto run = [&](auto ftag) {
    constexpr int F = decltype(ftag)::value;
    if (fx)
      hipLaunchKernelGGL((k_refl_finish<F, true, true>), grid, dim3(kBlock), 0, s, sc, levels, n_levels, band_first,
                         band_rays, out, ao, stats);
    else if (vis)
      hipLaunchKernelGGL((k_refl_finish<0, true, false>), grid, dim3(kBlock), 0, s, sc, levels, n_levels,
                         band_first, band_rays, out, ao, stats);
    else
      hipLaunchKernelGGL((k_refl_finish<0, false, false>), grid, dim3(kBlock), 0, s, sc, levels, n_levels,
                         band_first, band_rays, out, ao, stats);
  };
  if (!fx || formula_generic(formula)) run(std::integral_constant<int, 0>{});
  else if (formula == 7) run(std::integral_constant<int, 7>{});
  else if (formula == 8) run(std::integral_constant<int, 8>{});
  else run(std::integral_constant<int, 2>{});
  return hipGetLastError();
}

hipError_t launch_march(int precision, const DevScene *sc, const BandBufs &b, int band_first, int band_rays,
                        int formula, int plain, int persist_blocks, unsigned long long *stats, hipStream_t s,
                        bool ip8) {
  if (band_rays <= 0) return hipSuccess;
  // generic formulas: FP64, or the FP64 march of FP32_FIXUP (FP32 records)
  if (formula_generic(formula) && precision > 1) return hipErrorInvalidValue;
  const dim3 grid((unsigned)std::max(1, persist_blocks));
  auto run = [&](auto tag, auto rtag) {
    using T = decltype(tag);
    using RT = decltype(rtag);
    auto go = [&](auto plain_tag) {
      constexpr bool Pl = decltype(plain_tag)::value;
      if constexpr (Pl) {
        // the caller sets ip8 only where the trig-free iteration is wanted
        if (formula == 2 && ip8) {
          hipLaunchKernelGGL((k_march<T, RT, 2, true, 8>), grid, dim3(kBlock), 0, s, sc, b, band_first, band_rays,
                             stats);
          return;
        }
      }
      if (formula == 7)
        hipLaunchKernelGGL((k_march<T, RT, 7, Pl>), grid, dim3(kBlock), 0, s, sc, b, band_first, band_rays, stats);
      else if (formula == 8)
        hipLaunchKernelGGL((k_march<T, RT, 8, Pl>), grid, dim3(kBlock), 0, s, sc, b, band_first, band_rays, stats);
      else
        hipLaunchKernelGGL((k_march<T, RT, 2, Pl>), grid, dim3(kBlock), 0, s, sc, b, band_first, band_rays, stats);
    };
    if (formula_generic(formula)) {
      if constexpr (std::is_same<T, double>::value)
        hipLaunchKernelGGL((k_march<double, RT, 0, false>), grid, dim3(kBlock), 0, s, sc, b, band_first, band_rays,
                           stats);
      return;
    }
    if (plain) go(std::true_type{});
    else go(std::false_type{});
  };
  if (precision == 0) run(double{}, double{});
  else if (precision == 1) run(double{}, float{});
  else run(float{}, float{});
  return hipGetLastError();
}

hipError_t launch_march_fix(int precision, const DevScene *sc, const BandBufs &b, int band_first, int formula,
                            int persist_blocks, unsigned long long *stats, hipStream_t s, bool ip8) {
  const dim3 grid((unsigned)std::max(1, persist_blocks));
  if (formula_generic(formula) && precision > 1) return hipErrorInvalidValue;
  if (formula_generic(formula)) {
    if (precision == 0)
      hipLaunchKernelGGL((k_march_fix<double, 0>), grid, dim3(kBlock), 0, s, sc, b, band_first, stats);
    else
      hipLaunchKernelGGL((k_march_fix<float, 0>), grid, dim3(kBlock), 0, s, sc, b, band_first, stats);
    return hipGetLastError();
  }
  auto run = [&](auto rtag) {
    using RT = decltype(rtag);
    if (formula == 7)
      hipLaunchKernelGGL((k_march_fix<RT, 7>), grid, dim3(kBlock), 0, s, sc, b, band_first, stats);
    else if (formula == 8)
      hipLaunchKernelGGL((k_march_fix<RT, 8>), grid, dim3(kBlock), 0, s, sc, b, band_first, stats);
    else {
      bool done = false;
      if (ip8) {
        hipLaunchKernelGGL((k_march_fix<RT, 2, 8>), grid, dim3(kBlock), 0, s, sc, b, band_first, stats);
        done = true;
      }
      if (!done) hipLaunchKernelGGL((k_march_fix<RT, 2>), grid, dim3(kBlock), 0, s, sc, b, band_first, stats);
    }
    if (precision == 2) {
      if (formula == 7)
        hipLaunchKernelGGL((k_march_redo<RT, 7>), grid, dim3(kBlock), 0, s, sc, b, band_first, stats);
      else if (formula == 8)
        hipLaunchKernelGGL((k_march_redo<RT, 8>), grid, dim3(kBlock), 0, s, sc, b, band_first, stats);
      else
        hipLaunchKernelGGL((k_march_redo<RT, 2>), grid, dim3(kBlock), 0, s, sc, b, band_first, stats);
    }
  };
  if (precision == 0) run(double{});
  else run(float{});
  return hipGetLastError();
}

__global__ void k_band_reset(unsigned long long *counters) {
  if (threadIdx.x < CNT_COUNT) counters[threadIdx.x] = 0;
}

hipError_t launch_band_reset(const BandBufs &b, hipStream_t s) {
  hipLaunchKernelGGL(k_band_reset, dim3(1), dim3(64), 0, s, b.counters);
  return hipGetLastError();
}

__global__ void k_zero(unsigned long long *p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

hipError_t launch_zero(unsigned long long *p, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, s, p, n);
  return hipGetLastError();
}

hipError_t launch_shade(int precision, const DevScene *sc, const BandBufs &b, int band_first, AODev ao,
                        ShadeBufs sbuf, int persist_blocks, unsigned long long *stats, hipStream_t s,
                        hipEvent_t ev_ao0, hipEvent_t ev_ao1) {
  const int formula = sbuf.formula;
  // persistent grids: enough workgroups to fill every CU several times; the
  // kernels loop over the band's hits
  const dim3 grid((unsigned)std::max(1, persist_blocks));
  auto run = [&](auto tag) -> hipError_t {
    using T = decltype(tag);
    SurfBufs<T> sb;
    sb.base = (T *)sbuf.base;
    sb.spec = (T *)sbuf.spec;
    sb.col = (T *)sbuf.col;
    sb.fast_ao = (T *)sbuf.fast_ao;
    sb.ao = (T *)sbuf.ao;
    sb.aux = (T *)sbuf.aux;
    sb.n = sbuf.n;
    const bool generic = formula_generic(formula);
    auto surf = [&](auto deep_tag) {
      constexpr bool D = decltype(deep_tag)::value;
      if constexpr (!D) {
        if (generic) {
          hipLaunchKernelGGL((k_surface<T, 0, false>), grid, dim3(kBlock), 0, s, sc, b, band_first, sb, stats);
          return;
        }
      }
      if (formula == 7)
        hipLaunchKernelGGL((k_surface<T, 7, D>), grid, dim3(kBlock), 0, s, sc, b, band_first, sb, stats);
      else if (formula == 8)
        hipLaunchKernelGGL((k_surface<T, 8, D>), grid, dim3(kBlock), 0, s, sc, b, band_first, sb, stats);
      else
        hipLaunchKernelGGL((k_surface<T, 2, D>), grid, dim3(kBlock), 0, s, sc, b, band_first, sb, stats);
    };
    auto aok = [&](auto deep_tag, auto plain_tag) {
      constexpr bool D = decltype(deep_tag)::value;
      constexpr bool Pl = decltype(plain_tag)::value;
      if constexpr (!D && Pl && std::is_same<T, float>::value) {
        if (formula == 7)
          hipLaunchKernelGGL((k_ao_f32_plain<T, 7>), grid, dim3(kBlock), 0, s, sc, b, ao, sb, stats);
        else if (formula == 8)
          hipLaunchKernelGGL((k_ao_f32_plain<T, 8>), grid, dim3(kBlock), 0, s, sc, b, ao, sb, stats);
        else if (sbuf.ip8)
          hipLaunchKernelGGL((k_ao_f32_plain<T, 2, 8>), grid, dim3(kBlock), 0, s, sc, b, ao, sb, stats);
        else
          hipLaunchKernelGGL((k_ao_f32_plain<T, 2>), grid, dim3(kBlock), 0, s, sc, b, ao, sb, stats);
        return;
      }
      if constexpr (!D && !Pl) {
        if (generic) {
          hipLaunchKernelGGL((k_ao<T, 0, false, false>), grid, dim3(kBlock), 0, s, sc, b, ao, sb, stats);
          return;
        }
      }
      if constexpr (!D && Pl && std::is_same<T, double>::value) {
        if (formula == 2 && sbuf.ip8) {  // FP64_DEMOIVRE
          hipLaunchKernelGGL((k_ao<T, 2, false, true, 8>), grid, dim3(kBlock), 0, s, sc, b, ao, sb, stats);
          return;
        }
      }
      if (formula == 7)
        hipLaunchKernelGGL((k_ao<T, 7, D, Pl>), grid, dim3(kBlock), 0, s, sc, b, ao, sb, stats);
      else if (formula == 8)
        hipLaunchKernelGGL((k_ao<T, 8, D, Pl>), grid, dim3(kBlock), 0, s, sc, b, ao, sb, stats);
      else
        hipLaunchKernelGGL((k_ao<T, 2, D, Pl>), grid, dim3(kBlock), 0, s, sc, b, ao, sb, stats);
    };
    // deep-zoom variants only for FP32 shading of scenes that need them
    const bool deep = sbuf.deep && sizeof(T) < 8 && !generic;
    if (deep) surf(std::true_type{});
    else surf(std::false_type{});
    if (sbuf.lights) {
      if (generic) {
        if constexpr (std::is_same<T, double>::value)
          hipLaunchKernelGGL((k_aux<T, 0, false>), grid, dim3(kBlock), 0, s, sc, b, band_first, sb, stats);
      } else {
        auto aux = [&](auto deep_tag) {
          constexpr bool D = decltype(deep_tag)::value;
          if (formula == 7)
            hipLaunchKernelGGL((k_aux<T, 7, D>), grid, dim3(kBlock), 0, s, sc, b, band_first, sb, stats);
          else if (formula == 8)
            hipLaunchKernelGGL((k_aux<T, 8, D>), grid, dim3(kBlock), 0, s, sc, b, band_first, sb, stats);
          else
            hipLaunchKernelGGL((k_aux<T, 2, D>), grid, dim3(kBlock), 0, s, sc, b, band_first, sb, stats);
        };
        if (deep) aux(std::true_type{});
        else aux(std::false_type{});
      }
    }
    if (ev_ao0) {
      hipError_t e = hipEventRecord(ev_ao0, s);
      if (e != hipSuccess) return e;
    }
    if (sbuf.slow_ao) {
      // plain scenes (no primitives, limits, interior mode, iteration
      // threshold, pre-folds or Julia constant) take a variant without those
      // branches and the scene constants they keep live
      if (deep) aok(std::true_type{}, std::false_type{});
      else if (sbuf.plain && !generic) aok(std::false_type{}, std::true_type{});
      else aok(std::false_type{}, std::false_type{});
    }
    if (ev_ao1) {
      hipError_t e = hipEventRecord(ev_ao1, s);
      if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_compose<T, T>), grid, dim3(kBlock), 0, s, sc, b, ao, sb);
    return hipGetLastError();
  };
  // generic formulas in FP32: FP32_FIXUP without aux lights (the host keeps
  // other generic scenes in FP64)
  if (formula_generic(formula) && (precision > 1 || (precision == 1 && sbuf.lights))) return hipErrorInvalidValue;
  if (precision == 0) return run(double{});
  return run(float{});
}

hipError_t launch_finish(int precision, const DevScene *sc, const BandBufs &b, int band_first, int band_rays,
                         FrameOut out, hipStream_t s, int formula, int fx, int vis, int persist_blocks,
                         unsigned long long *stats, AODev ao) {
  if (band_rays <= 0) return hipSuccess;
  if (fx) {
    if (formula_generic(formula) && precision != 0) return hipErrorInvalidValue;
    const dim3 grid((unsigned)std::max(1, std::min(persist_blocks, (band_rays + kBlock - 1) / kBlock)));
    if (formula_generic(formula)) {
      hipLaunchKernelGGL((k_finish_fx<double, double, 0>), grid, dim3(kBlock), 0, s, sc, b, band_first, band_rays,
                         out, stats, ao);
      return hipGetLastError();
    }
    auto run = [&](auto tag, auto rtag) {
      using T = decltype(tag);
      using RT = decltype(rtag);
      if (formula == 7)
        hipLaunchKernelGGL((k_finish_fx<T, RT, 7>), grid, dim3(kBlock), 0, s, sc, b, band_first, band_rays, out,
                           stats, ao);
      else if (formula == 8)
        hipLaunchKernelGGL((k_finish_fx<T, RT, 8>), grid, dim3(kBlock), 0, s, sc, b, band_first, band_rays, out,
                           stats, ao);
      else
        hipLaunchKernelGGL((k_finish_fx<T, RT, 2>), grid, dim3(kBlock), 0, s, sc, b, band_first, band_rays, out,
                           stats, ao);
    };
    if (precision == 0) run(double{}, double{});
    else run(float{}, float{});
    return hipGetLastError();
  }
  auto plain = [&](auto tag, auto vis_tag) {
    using T = decltype(tag);
    constexpr bool V = decltype(vis_tag)::value;
    hipLaunchKernelGGL((k_finish<T, T, V>), grid_for(band_rays), dim3(kBlock), 0, s, sc, b, band_first, band_rays,
                       out);
  };
  if (precision == 0) {
    if (vis) plain(double{}, std::true_type{});
    else plain(double{}, std::false_type{});
  } else {
    if (vis) plain(float{}, std::true_type{});
    else plain(float{}, std::false_type{});
  }
  return hipGetLastError();
}

}  // namespace mbr
