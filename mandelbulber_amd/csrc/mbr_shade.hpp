// mbr_shade.hpp — shaders of the render core, templated on arithmetic type.
//
//   normals<T>            CalculateNormals       src/shaders.cpp:675-757
//   main_shadow<T>        MainShadow             src/shaders.cpp:480-570
//   fast_ao2<T>           FastAmbientOcclusion2  src/shaders.cpp:585-605
//   ambient_occlusion<T>  AmbientOcclusion       src/shaders.cpp:607-673
//   surface_colour<T>     SurfaceColour          src/shaders.cpp:813-926
//   index_to_colour       IndexToColour          src/shaders.cpp:1273-1305
//   object_shader<T>      ObjectShader           src/shaders.cpp:21-112
//   background<T>         BackgroundShader       src/shaders.cpp:114-183
//   volumetric<T>         VolumetricShader       src/shaders.cpp:185-478
//   aux_shadow<T>         AuxShadow              src/shaders.cpp:1075-1132
//   aux_lights<T>         AuxLightsShader        src/shaders.cpp:1051-1073
//                         + LightShading         src/shaders.cpp:997-1049
//
// The volumetric lights' per-step shadow marches run in k_finish_fx through
// the `fx` hook of volumetric<T>.  Features outside the render core (env
// mapping, fake lights, iteration fog) are rejected at mbr_set_scene, so
// their terms are zero here.
#pragma once

#include "mbr_fractal.hpp"

namespace mbr {

// Upper bound on the DE steps of one shadow or AO ray.  The reference loops
// have no bound (src/shaders.cpp:508, :625); a ray that reaches this many
// steps without terminating is ended as unoccluded, so every lane finishes.
constexpr int kMaxShadeSteps = 20000;

struct LaneCount {
  uint32_t de_calls;
  uint32_t iterations;
  MBR_HD void add(const DEInfo &i) {
    de_calls += 1;
    iterations += (uint32_t)i.work;
  }
};

template <class T>
struct ShaderIn {
  V3<T> point, view, normal, light;
  V3<double> point64;  // the hit point in FP64 (distance estimates start from it)
  T dist_thresh, lastDist, delta, depth;
  T detail;  // calcParam->doubles.detailSize as the reference mutates it
  int object;
};

template <class T>
struct RGBT {
  T R, G, B;
};

// Distance estimate of the shaders.  F is the formula when the caller was
// compiled for one (a kernel per formula), 0 for the kernels of the generic
// FP64 formulas.
template <class T, int F = 0>
MBR_HD T de(const DevScene &sc, V3<T> p, T detail, int N, int minN, bool normalMode, DEInfo &info,
            LaneCount &cnt) {
  T d;
  if constexpr (F == 0) d = calc_distance_generic<T>(sc.frac<T>(), p, detail, N, minN, normalMode, info);
  else d = calc_distance_f<T, F>(sc.frac<T>(), p, detail, N, minN, normalMode, info);
  cnt.add(info);
  return d;
}

// Distance estimate at the hit point plus an offset.  Deep (FP32 callers
// of deep-zoom scenes): the point is formed in FP64 and the estimate starts
// in FP64 (calc_distance_p64); otherwise the point and the estimate are in
// the caller's precision.
template <class T, int F, bool Deep>
MBR_HD T de_at(const DevScene &sc, const ShaderIn<T> &in, V3<T> off, T detail, int N, int minN, bool normalMode,
               DEInfo &info, LaneCount &cnt) {
  if constexpr (Deep && sizeof(T) < 8 && F != 0) {
    const T d = calc_distance_p64<T, F>(sc.fd, sc.frac<T>(), in.point64 + cvt<double>(off), detail, N, minN,
                                        normalMode, info);
    cnt.add(info);
    return d;
  } else {
    return de<T, F>(sc, in.point + off, detail, N, minN, normalMode, info, cnt);
  }
}

template <class T, int F = 0, bool Deep = false>
MBR_HD V3<T> normals(const DevScene &sc, ShaderIn<T> &in, LaneCount &cnt) {
  const ShadeK<T> &k = sc.shade<T>();
  V3<T> n = v3<T>(0, 0, 0);
  DEInfo info;
  const int N5 = sc.N * 5;
  if constexpr (F == 0 && sizeof(T) < 8) {
    // FP32 shading of the generic formulas: the four estimates in FP64 at
    // the FP64 hit point (differences of FP32 estimates a stencil step
    // apart keep few digits)
    if (!sc.slow_shading) {
      T delta = in.delta * k.smoothness;
      if (sc.interior) delta = in.dist_thresh * k.quality * T(0.2) * k.smoothness;
      const double dd = (double)delta, det = (double)in.detail;
      auto de64 = [&](double ox, double oy, double oz) {
        const double d = calc_distance_generic<double>(sc.fd, in.point64 + v3<double>(ox, oy, oz), det, N5, 0, true,
                                                       info);
        cnt.add(info);
        return d;
      };
      const double s1 = de64(0, 0, 0);
      n = v3<T>((T)(de64(dd, 0, 0) - s1), (T)(de64(0, dd, 0) - s1), (T)(de64(0, 0, dd) - s1));
      if (n.x == 0 && n.y == 0 && n.z == 0) n.x = 1;
      else n = normalize(n);
      return n;
    }
  }
  if (!sc.slow_shading) {
    T delta = in.delta * k.smoothness;
    if (sc.interior) delta = in.dist_thresh * k.quality * T(0.2) * k.smoothness;
    T s1 = de_at<T, F, Deep>(sc, in, v3<T>(0, 0, 0), in.detail, N5, 0, true, info, cnt);
    T s2 = de_at<T, F, Deep>(sc, in, v3<T>(delta, 0, 0), in.detail, N5, 0, true, info, cnt);
    T s3 = de_at<T, F, Deep>(sc, in, v3<T>(0, delta, 0), in.detail, N5, 0, true, info, cnt);
    T s4 = de_at<T, F, Deep>(sc, in, v3<T>(0, 0, delta), in.detail, N5, 0, true, info, cnt);
    n = v3<T>(s2 - s1, s3 - s1, s4 - s1);
  } else {
    T delta = in.delta * k.smoothness;
    if (sc.interior) delta = in.dist_thresh * k.quality * T(0.2) * k.smoothness;
    // the 11^3 stencil with the reference's accumulated loop variables
    for (double px = -1.0; px <= 1.0; px += 0.2)
      for (double py = -1.0; py <= 1.0; py += 0.2)
        for (double pz = -1.0; pz <= 1.0; pz += 0.2) {
          V3<T> p2 = v3<T>((T)px, (T)py, (T)pz);
          T d = de_at<T, F, Deep>(sc, in, p2 * delta, in.detail, N5, sc.minN, true, info, cnt);
          n = n + p2 * d;
        }
  }
  if (n.x == 0 && n.y == 0 && n.z == 0) n.x = 1;
  else n = normalize(n);
  return n;
}

// Threshold of a shadow-march point when volumetric lights or iteration fog
// are on (src/shaders.cpp:511-519, :1102-1110): from its distance to the eye.
template <class T>
MBR_HD float fx_thresh(const DevScene &sc, V3<T> point2) {
  const ShadeK<T> &k = sc.shade<T>();
  if (!sc.constant_thresh) {
    const V3<T> eye = v3<T>((T)sc.start[0], (T)sc.start[1], (T)sc.start[2]);
    return (float)(length(point2 - eye) * k.resolution * k.persp / k.quality);
  }
  return (float)k.quality;
}

// IterOpacity (src/shaders.cpp:1214-1222)
template <class T>
MBR_HD T iter_opacity(T step, T iters, T maxN, T trim, T opacitySp) {
  T opacity = (iters - trim) / maxN;
  if (opacity < T(0)) opacity = 0;
  opacity *= opacity;
  opacity *= step * opacitySp;
  if (opacity > T(1)) opacity = 1;
  return opacity;
}

// the fog opacity a shadow march collects at an estimate (0 without
// iteration fog)
template <class T>
MBR_HD T fog_opacity(const DevScene &sc, T dist_step, int iters) {
  if (!sc.iter_fog) return T(0);
  return iter_opacity<T>(dist_step, T(iters), T(sc.N), (T)sc.iter_fog_trim, (T)sc.iter_fog_opacity);
}

template <class T, int F = 0, bool Deep = false>
MBR_HD T main_shadow(const DevScene &sc, ShaderIn<T> &in, LaneCount &cnt) {
  const ShadeK<T> &k = sc.shade<T>();
  T factor = in.delta / k.resolution;
  if (!sc.penetrating) factor = k.view_distance_max;
  T dist = in.dist_thresh;
  in.detail = in.dist_thresh;
  const T DE_factor = sc.shadow_fx ? T(1) : k.DE_factor;
  T start = in.dist_thresh;
  if (sc.interior) start = in.dist_thresh * DE_factor;
  T shadowTemp = 1;
  T maxSoft = 0;
  DEInfo info;
  int guard = 0;
  for (T i = start; i < factor && guard < kMaxShadeSteps; i += dist * DE_factor, guard++) {

    const float dist_thresh = sc.shadow_fx ? fx_thresh<T>(sc, in.point + in.light * i) : (float)in.dist_thresh;
    in.detail = (T)dist_thresh;
    dist = de_at<T, F, Deep>(sc, in, in.light * i, in.detail, sc.N, sc.minN, false, info, cnt);
    if (sc.bsoft) {
      T angle = (dist - (T)dist_thresh) / i;
      if (angle < 0) angle = 0;
      if (dist < (T)dist_thresh) angle = 0;
      T softShadow = (T(1) - angle / k.soft_range);
      if (sc.penetrating) softShadow *= (factor - i) / factor;
      if (softShadow < 0) softShadow = 0;
      if (softShadow > maxSoft) maxSoft = softShadow;
    }
    if (sc.iter_fog) shadowTemp -= fog_opacity<T>(sc, dist * DE_factor, info.iters) * (factor - i) / factor;
    if (dist < (T)dist_thresh || shadowTemp < T(0)) {
      shadowTemp -= (factor - i) / factor;
      if (!sc.penetrating) shadowTemp = 0;
      if (shadowTemp < T(0)) shadowTemp = 0;
      break;
    }
  }
  return sc.bsoft ? (T(1) - maxSoft) : shadowTemp;
}

template <class T, int F = 0, bool Deep = false>
MBR_HD T fast_ao2(const DevScene &sc, ShaderIn<T> &in, LaneCount &cnt) {
  const ShadeK<T> &k = sc.shade<T>();
  T delta = in.dist_thresh;
  T aoTemp = 0;
  const int q = sc.ao_quality;
  DEInfo info;
  T inv2i = T(1);
  for (int i = 1; i < q * q; i++) {
    inv2i *= T(0.5);  // 1/pow(2,i), exact
    T scan = T(i * i) * delta;
    T dist = de_at<T, F, Deep>(sc, in, in.normal * scan, in.detail, sc.N, sc.minN, false, info, cnt);
    aoTemp += inv2i * (scan - k.fast_ao_tune * dist) / in.dist_thresh;
  }
  T ao = T(1) - T(0.2) * aoTemp;
  if (ao < 0) ao = 0;
  return ao;
}

template <class T>
struct AOTable {
  const T *dir;      // count x 4 (x, y, z, unused)
  const int *rgb;    // count x 4 (R, G, B, unused)
  int count;
};

template <class T, int F = 0, bool Deep = false>
MBR_HD RGBT<T> ambient_occlusion(const DevScene &sc, ShaderIn<T> &in, const AOTable<T> &tab, LaneCount &cnt) {
  const ShadeK<T> &k = sc.shade<T>();
  RGBT<T> ao = {0, 0, 0};
  const T start_dist = in.delta;
  const T end_dist = in.delta / k.resolution;
  const float dist_thresh = (float)in.dist_thresh;
  DEInfo info;
  for (int v = 0; v < tab.count; v++) {
    const V3<T> dir = v3<T>(tab.dir[4 * v + 0], tab.dir[4 * v + 1], tab.dir[4 * v + 2]);
    T dist = in.lastDist;
    T shadowTemp = 1;
    int guard = 0;
    for (T r = start_dist; r < end_dist && guard < kMaxShadeSteps; r += dist * T(2), guard++) {
      dist = de_at<T, F, Deep>(sc, in, dir * r, in.detail, sc.N, sc.minN, false, info, cnt);
      if (dist < (T)dist_thresh || info.max_iter || shadowTemp < T(0)) {
        shadowTemp -= (end_dist - r) / end_dist;
        if (shadowTemp < T(0)) shadowTemp = 0;
        break;
      }
    }
    ao.R += shadowTemp * T(tab.rgb[4 * v + 0]);
    ao.G += shadowTemp * T(tab.rgb[4 * v + 1]);
    ao.B += shadowTemp * T(tab.rgb[4 * v + 2]);
  }
  const T den = T(tab.count) * T(256);
  ao.R /= den;
  ao.G /= den;
  ao.B /= den;
  return ao;
}

// AuxShadow: march from the point towards the light until the light's
// distance.
template <class T, int F = 0, bool Deep = false>
MBR_HD T aux_shadow(const DevScene &sc, ShaderIn<T> &in, T distance, V3<T> lightVector, LaneCount &cnt) {
  const ShadeK<T> &k = sc.shade<T>();
  T dist = in.delta;
  T shadowTemp = 1;
  const T DE_factor = sc.shadow_fx ? T(1) : k.DE_factor;
  DEInfo info;
  int guard = 0;
  for (T i = in.delta; i < distance && guard < kMaxShadeSteps; i += dist * DE_factor, guard++) {
    dist = de_at<T, F, Deep>(sc, in, lightVector * i, in.detail, sc.N, sc.minN, false, info, cnt);
    if (sc.iter_fog) shadowTemp -= fog_opacity<T>(sc, dist * DE_factor, info.iters) * (distance - i) / distance;
    const float dist_thresh = sc.shadow_fx ? fx_thresh<T>(sc, in.point + lightVector * i) : (float)in.dist_thresh;
    if (dist < (T)dist_thresh || info.max_iter || shadowTemp < T(0)) {
      if (sc.penetrating) {
        shadowTemp -= (distance - i) / distance;
        if (shadowTemp < T(0)) shadowTemp = 0;
      } else {
        shadowTemp = 0;
      }
      break;
    }
  }
  return shadowTemp;
}

// AuxLightsShader: the sum over the lights that take part of LightShading's
// diffuse (`shade`) and specular (`spec`) terms, in the lights' order.
template <class T, int F = 0, bool Deep = false>
MBR_HD void aux_lights(const DevScene &sc, ShaderIn<T> &in, RGBT<T> &shade, RGBT<T> &spec, LaneCount &cnt) {
  const ShadeK<T> &k = sc.shade<T>();
  const T *tab = sc.lights<T>();
  shade = RGBT<T>{0, 0, 0};
  spec = RGBT<T>{0, 0, 0};
  for (int l = 0; l < sc.n_lights; l++) {
    const T *L = tab + 8 * l;
    const V3<T> d = v3<T>(L[0], L[1], L[2]) - in.point;
    const T distance = length(d);
    const V3<T> lightVector = normalize(d);
    const T intensity = k.aux_intensity * T(100) * L[3] / (distance * distance) / T(sc.light_div);
    T sh = dot(in.normal, lightVector);
    if (sh < T(0)) sh = 0;
    sh = sh * intensity;
    if (sh > T(500)) sh = 500;
    const V3<T> half = normalize(lightVector - in.view);
    T sh2 = dot(in.normal, half);
    if (sh2 < T(0)) sh2 = 0;
    sh2 = Math<T>::pow(sh2, T(30));
    sh2 *= intensity * k.specular;
    if (sh2 > T(15)) sh2 = 15;
    if ((sh > T(0.01) || sh2 > T(0.01)) && sc.shadow) {
      const T light = aux_shadow<T, F, Deep>(sc, in, distance, lightVector, cnt);
      sh *= light;
      sh2 *= light;
    } else if (sc.shadow) {
      sh = 0;
      sh2 = 0;
    }
    shade.R += sh * L[4] / T(65536);
    shade.G += sh * L[5] / T(65536);
    shade.B += sh * L[6] / T(65536);
    spec.R += sh2 * L[4] / T(65536);
    spec.G += sh2 * L[5] / T(65536);
    spec.B += sh2 * L[6] / T(65536);
  }
}

// FakeLights (src/shaders.cpp:1225-1270): the orbit-trap sum at the hit and
// one step of dist_thresh x smoothness along each axis give the fake light's
// direction; Lambert and a power-30 specular term of intensity / sum.
// Fake-light scenes render in FP64.
template <class T>
MBR_HD void fake_lights(const DevScene &sc, const ShaderIn<T> &in, RGBT<T> &shade, RGBT<T> &spec, LaneCount &cnt) {
  const FractalK<T> &f = sc.frac<T>();
  const T delta = in.dist_thresh * sc.shade<T>().smoothness;
  int L = 0;
  const T rr = compute<T, MODE_ORBIT_TRAP>(f, in.point, sc.N, &L);
  const T fakeLight = T(sc.fl_intensity) / rr;
  const T r = T(1) / (rr + T(1e-30));
  const T rx = T(1) / (compute<T, MODE_ORBIT_TRAP>(f, in.point + v3<T>(delta, 0, 0), sc.N, &L) + T(1e-30));
  const T ry = T(1) / (compute<T, MODE_ORBIT_TRAP>(f, in.point + v3<T>(0, delta, 0), sc.N, &L) + T(1e-30));
  const T rz = T(1) / (compute<T, MODE_ORBIT_TRAP>(f, in.point + v3<T>(0, 0, delta), sc.N, &L) + T(1e-30));
  cnt.de_calls += 4;
  V3<T> n = v3<T>(r - rx, r - ry, r - rz);
  if (n.x == 0 && n.y == 0 && n.z == 0) n.x = 0;
  else n = normalize(n);
  T lam = fakeLight * dot(in.normal, n);
  if (lam < 0) lam = 0;
  shade = RGBT<T>{lam, lam, lam};
  const V3<T> half = normalize(n - in.view);
  T fs = dot(in.normal, half);
  if (fs < T(0)) fs = 0;
  fs = Math<T>::pow(fs, T(30)) * fakeLight;
  if (fs > T(15)) fs = 15;
  spec = RGBT<T>{fs, fs, fs};
}

MBR_HD void index_to_colour(int index, const int (*palette)[3], int out[3]) {
  if (index < 0) {
    out[0] = palette[255][0]; out[1] = palette[255][1]; out[2] = palette[255][2];
    return;
  }
  out[0] = out[1] = out[2] = 255;
  index = index % (255 * 256);
  int kol = index / 256;
  if (kol < 255) {
    int delta = index % 256;
    for (int c = 0; c < 3; c++) {
      double c1 = palette[kol][c], c2 = palette[kol + 1][c];
      double ck = (c2 - c1) / 256.0;
      out[c] = (int)(c1 + ck * delta);
    }
  }
}

// G: the colouring orbit may be a generic formula's (compute's G)
template <class T, bool Deep = false, bool G = (sizeof(T) == 8)>
MBR_HD RGBT<T> surface_colour(const DevScene &sc, const ShaderIn<T> &in, LaneCount &cnt) {
  RGBT<T> out;
  const int *pc;
  switch (in.object) {
    case OBJ_FRACTAL: {
      int L;
      const int N10 = sc.N * 10;
      T cval;
      if (Deep && sizeof(T) < 8 && (double)in.detail < 2.0 * kMixedSwitch * length(in.point64))
        // deep zoom: the FP32 point would merge neighbouring pixels' orbits
        cval = (T)compute<double, MODE_COLOURING>(sc.fd, in.point64, N10, &L);
      else
        cval = compute<T, MODE_COLOURING, G>(sc.frac<T>(), in.point, N10, &L);
      cnt.de_calls += 1;
      cnt.iterations += (uint32_t)((L < N10) ? L + 1 : N10);
      int nrCol = x86_d2i(floor((double)cval));
      nrCol = x86_abs(nrCol) % (248 * 256);
      int colour[3] = {256, 256, 256};
      if (sc.coloring) {
        const ShadeK<T> &k = sc.shade<T>();
        int color_number;
        if (nrCol >= 248 * 256) color_number = nrCol;
        else color_number = x86_d2i((double)(T(nrCol) * k.coloring_speed + T(256) * k.palette_offset)) % 65536;
        index_to_colour(color_number, sc.palette, colour);
      }
      out.R = T(colour[0]) / T(256); out.G = T(colour[1]) / T(256); out.B = T(colour[2]) / T(256);
      return out;
    }
    case OBJ_WATER: pc = sc.water_col; break;
    case OBJ_BOX: pc = sc.box_col; break;
    case OBJ_BOX_INV: pc = sc.ibox_col; break;
    case OBJ_SPHERE: pc = sc.sphere_col; break;
    case OBJ_SPHERE_INV: pc = sc.isphere_col; break;
    default: pc = sc.plane_col; break;
  }
  out.R = T(pc[0]) / T(65536); out.G = T(pc[1]) / T(65536); out.B = T(pc[2]) / T(65536);
  return out;
}

// ObjectShader: returns object + specular terms; *colour_out is the surface
// colour (colourBuffer source).
template <class T, int F = 0, bool Deep = false>
MBR_HD RGBT<T> object_shader(const DevScene &sc, ShaderIn<T> &in, const AOTable<T> &tab, RGBT<T> &colour_out,
                             LaneCount &cnt) {
  const ShadeK<T> &k = sc.shade<T>();
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
  colour_out = colour;
  RGBT<T> ambient = {0, 0, 0};
  if (sc.ao) {
    if (sc.fast_ao) {
      T a = fast_ao2<T, F, Deep>(sc, in, cnt);
      ambient = RGBT<T>{a, a, a};
    } else {
      ambient = ambient_occlusion<T, F, Deep>(sc, in, tab, cnt);
    }
  }
  RGBT<T> out;
  out.R = (ambient.R * k.global_ilum + k.main_light[0] * shade * shadow) * colour.R + k.main_light[0] * specular * shadow;
  out.G = (ambient.G * k.global_ilum + k.main_light[1] * shade * shadow) * colour.G + k.main_light[1] * specular * shadow;
  out.B = (ambient.B * k.global_ilum + k.main_light[2] * shade * shadow) * colour.B + k.main_light[2] * specular * shadow;
  return out;
}

// cTexture::Pixel / Interpolation (src/texture.cpp:83-111) on 8-bit RGB
// rows: bilinear over the texels (ix, iy), (ix + 1, iy), (ix, iy + 1),
// (ix + 1, iy + 1), each channel's sum truncated to 8 bits; black outside
// [0, w) x [0, h - 1).  ix + 1 at the last column reads the next row's
// first texel, as the reference's flat index does; the one texel past the
// bitmap's end that it then reads (last column, second-last row) reads
// black here.
MBR_HD inline void tex_pixel(const uint8_t *bm, int w, int h, double x, double y, int out[3]) {
#pragma clang fp contract(off)
  out[0] = out[1] = out[2] = 0;
  if (!(x >= 0 && x < w && y >= 0 && y < h - 1.0)) return;
  const int ix = (int)x, iy = (int)y;
  const double rx = x - (int)x, ry = y - (int)y;
  const long long n = (long long)w * h;
  const long long i1 = (long long)iy * w + ix, i2 = i1 + 1;
  const long long i3 = (long long)(iy + 1) * w + ix, i4 = i3 + 1;
  for (int c = 0; c < 3; c++) {
    const double k1 = bm[3 * i1 + c], k2 = bm[3 * i2 + c], k3 = bm[3 * i3 + c];
    const double k4 = i4 < n ? bm[3 * i4 + c] : 0;
    out[c] = (int)(uint8_t)(int)(k1 * (1.0 - rx) * (1.0 - ry) + k2 * rx * (1.0 - ry) + k3 * (1.0 - rx) * ry +
                                 k4 * (rx * ry));
  }
}

// BackgroundShader's texture branches (src/shaders.cpp:118-157): the view
// vector's angles index the background texture, as a fisheye dome of two
// hemispheres side by side or as an equirectangular map.
MBR_HD inline void background_texture(const DevScene &sc, V3<double> v, double out[3]) {
#pragma clang fp contract(off)
  const double pi = 3.14159265358979323846;
  double alfa = atan2(v.y, v.x);
  double beta = atan2(v.z, sqrt(v.x * v.x + v.y * v.y));
  double texX, texY;
  if (sc.fulldome) {
    const int texWidth = (int)(sc.bg_w * 0.5);
    const int texHeight = sc.bg_h;
    int offset = 0;
    if (beta < 0) {
      beta = -beta;
      alfa += pi;
      offset = texWidth;
    }
    texX = 0.5 * texWidth + cos(alfa) * (1.0 - beta / (0.5 * pi)) * texWidth * 0.5 + offset;
    texY = 0.5 * texHeight + sin(alfa) * (1.0 - beta / (0.5 * pi)) * texHeight * 0.5;
  } else {
    alfa = fmod(alfa + 2.5 * pi, 2 * pi);
    if (beta > 0.5 * pi) beta = 0.5 * pi - beta;
    if (beta < -0.5 * pi) beta = -0.5 * pi + beta;
    texX = alfa / (2.0 * pi) * sc.bg_w;
    texY = (beta / (pi) + 0.5) * sc.bg_h;
  }
  int px[3];
  tex_pixel(sc.bg_tex, sc.bg_w, sc.bg_h, texX, texY, px);
  for (int c = 0; c < 3; c++) out[c] = px[c] / 256.0;
}

// EnvMapping (src/shaders.cpp:785-812): the view vector mirrored about the
// normal indexes the environment texture by its angles (wrapped with the
// reference's 8.25 and 8 texture offsets).
MBR_HD inline void env_mapping(const DevScene &sc, V3<double> n, V3<double> view, double out[3]) {
#pragma clang fp contract(off)
  const double pi = 3.14159265358979323846;
  const double d = -(view.x * n.x + view.y * n.y + view.z * n.z);
  const double rx = n.x * 2.0 * d + view.x, ry = n.y * 2.0 * d + view.y, rz = n.z * 2.0 * d + view.z;
  const double alfa = atan2(ry, rx) + pi;
  double beta = atan2(rz, sqrt(rx * rx + ry * ry));
  const double w = sc.env_w, h = sc.env_h;
  if (beta > 0.5 * pi) beta = 0.5 * pi - beta;
  if (beta < -0.5 * pi) beta = -0.5 * pi + beta;
  double dtx = (alfa / (2.0 * pi)) * w + w * 8.25;
  double dty = (beta / (pi) + 0.5) * h + h * 8.0;
  dtx = fmod(dtx, w);
  dty = fmod(dty, h);
  if (dtx < 0) dtx = 0;
  if (dty < 0) dty = 0;
  int px[3];
  tex_pixel(sc.env_tex, sc.env_w, sc.env_h, dtx, dty, px);
  for (int c = 0; c < 3; c++) out[c] = px[c] / 256.0;
}

template <class T>
MBR_HD RGBT<T> background(const DevScene &sc, V3<T> view) {
  if (sc.textured_bg) {
    double t[3];
    background_texture(sc, v3<double>((double)view.x, (double)view.y, (double)view.z), t);
    return RGBT<T>{T(t[0]), T(t[1]), T(t[2])};
  }
  T grad = -view.z + T(1);  // viewVector . (0,0,-1) + 1
  const int *c_lo, *c_hi;
  if (grad < T(1)) {
    c_lo = sc.bg3; c_hi = sc.bg2;
  } else {
    grad = grad - T(1);
    c_lo = sc.bg2; c_hi = sc.bg1;
  }
  const T Ngrad = T(1) - grad;
  RGBT<T> out;
  // sRGB16 storage of the gradient before scaling
  out.R = T(x86_d2u16((double)(T(c_lo[0]) * Ngrad + T(c_hi[0]) * grad))) / T(65536);
  out.G = T(x86_d2u16((double)(T(c_lo[1]) * Ngrad + T(c_hi[1]) * grad))) / T(65536);
  out.B = T(x86_d2u16((double)(T(c_lo[2]) * Ngrad + T(c_hi[2]) * grad))) / T(65536);
  return out;
}

// Step record of one ray, read back to front.  rec(i) yields
// {step, distance, distThresh} of march step i.
template <class T>
struct VolOut {
  RGBT<T> rgb;
  T alpha;
  T opacity;
};

// The ray of a pixel, for the glow of visible aux lights: a march point is
// start + dir * scan, scan being the sum of the recorded steps up to it
// (src/Render3D.cpp:96-139: step i is recorded before scan advances).
struct VisRay {
  V3<double> start, dir;
  double scan_end;  // scan at the last record
};

// Glow of the visible aux lights over one group of march steps ending at
// `point` (src/shaders.cpp:241-300), in FP64.
template <class T>
MBR_HD void visible_lights(const DevScene &sc, V3<double> point, V3<double> view, double step, VolOut<T> &o) {
  const ShadeK<double> &k = sc.sd;
  const double *tab = sc.lights_d;
  double lowestLightSize = 1e10, lowestLightDist = 1e10;
  for (int l = 0; l < sc.n_lights; l++) {
    const double *L = tab + 8 * l;
    const double lightDist = length(point - v3<double>(L[0], L[1], L[2]));
    const double r2 = lightDist / (L[3] * k.aux_intensity * k.aux_visibility);
    if (r2 < lowestLightSize) {
      lowestLightSize = r2;
      lowestLightDist = lightDist;
    }
  }
  int smallSteps = x86_d2i(10.0 * step / (lowestLightDist + 1.0e-20));
  if (smallSteps > 50) smallSteps = 50;
  int s0 = 0, s1 = 1;
  double step2 = step;
  if (smallSteps > 0) {
    s0 = 1;
    s1 = smallSteps + 1;
    step2 = step / (smallSteps + 1.0);
  }
  double R = (double)o.rgb.R, G = (double)o.rgb.G, B = (double)o.rgb.B, A = (double)o.alpha;
  for (int ss = s0; ss < s1; ss++) {
    const V3<double> point2 = point - (view * step2) * (double)ss;
    for (int l = 0; l < sc.n_lights; l++) {
      const double *L = tab + 8 * l;
      const double lightDist = length(point2 - v3<double>(L[0], L[1], L[2]));
      const double lightSize = L[3] * k.aux_intensity * k.aux_visibility;
      double r2 = lightDist / lightSize;
      if (r2 > 1.0) r2 = 1.0;
      const double bell = (cos(r2 * 3.14159265358979323846) + 1.0) / (r2 * r2 + 0.02) * 0.3;
      const double lightDensity = step2 * bell / lightSize;
      R += lightDensity * L[4] / 65536.0;
      G += lightDensity * L[5] / 65536.0;
      B += lightDensity * L[6] / 65536.0;
      A += lightDensity;
    }
  }
  o.rgb.R = (T)R;
  o.rgb.G = (T)G;
  o.rgb.B = (T)B;
  o.alpha = (T)A;
}

struct NoFx {
  template <class T>
  MBR_HD void operator()(int, int, V3<double>, T, T, VolOut<T> &, float &) const {}
};

// VolumetricShader over the step records.  With a ray (vr), a step group's
// point is known: the glow of visible aux lights is added when `vis`, and
// `fx(stage, index, point, threshold, step, out, totalOpacity)` adds the
// volumetric lights (stage 0, before the fogs) and the iteration fog
// (stage 1, after them).
template <class T, class RecFn, class FxFn = NoFx>
MBR_HD VolOut<T> volumetric(const DevScene &sc, T depth, int stepCount, RGBT<T> px, T alpha, RecFn rec,
                            const VisRay *vr = nullptr, bool vis = false, FxFn fx = NoFx{}) {
  const ShadeK<T> &k = sc.shade<T>();
  VolOut<T> o;
  o.rgb = px;
  o.alpha = alpha;
  o.opacity = 0;
  float totalOpacity = 0.0f;
  const T glow = T(stepCount) * k.glow_intensity / T(512) * k.DE_factor;
  T glowN = T(1) - glow;
  if (glowN < T(0)) glowN = 0;
  const T glowR = (T(sc.glow1[0]) * glowN + T(sc.glow2[0]) * glow) / T(65536);
  const T glowG = (T(sc.glow1[1]) * glowN + T(sc.glow2[1]) * glow) / T(65536);
  const T glowB = (T(sc.glow1[2]) * glowN + T(sc.glow2[2]) * glow) / T(65536);
  T totalStep = 0;
  double scan = vr ? vr->scan_end : 0.0;
  for (int index = stepCount - 1; index > 0; index--) {
    T step, distance, dist_thresh;
    rec(index, step, distance, dist_thresh);
    const double scan_i = scan;  // scan of this record's point
    scan -= (double)step;
    totalStep += step;
    if (totalStep < dist_thresh * k.quality) continue;
    step = totalStep;
    totalStep = 0;

    const T glowOpacity = glow * step / depth;
    o.rgb.R = glowOpacity * glowR + (T(1) - glowOpacity) * o.rgb.R;
    o.rgb.G = glowOpacity * glowG + (T(1) - glowOpacity) * o.rgb.G;
    o.rgb.B = glowOpacity * glowB + (T(1) - glowOpacity) * o.rgb.B;
    o.alpha += glowOpacity;

    V3<double> point = v3<double>(0, 0, 0);
    if (vr) {
      point = vr->start + vr->dir * scan_i;
      if (vis) visible_lights<T>(sc, point, vr->dir, (double)step, o);
      fx(0, index, point, dist_thresh, step, o, totalOpacity);
    }

    if (sc.post_fog) {
      T fd = step / k.fog_visibility;
      if (fd > T(1)) fd = 1;
      o.rgb.R = fd * T(sc.post_fog_colour[0]) / T(65536) + (T(1) - fd) * o.rgb.R;
      o.rgb.G = fd * T(sc.post_fog_colour[1]) / T(65536) + (T(1) - fd) * o.rgb.G;
      o.rgb.B = fd * T(sc.post_fog_colour[2]) / T(65536) + (T(1) - fd) * o.rgb.B;
      totalOpacity = (float)(fd + (T(1) - fd) * T(totalOpacity));
      o.alpha = fd + (T(1) - fd) * o.alpha;
    }

    if (k.fog_density > T(0)) {
      const T fogReduce = k.fog_reduce;
      const T densityTemp = (step * fogReduce) / (distance * distance + fogReduce * fogReduce);
      T kk = distance / k.fog_c1_dist;
      if (kk > T(1)) kk = 1;
      T kn = T(1) - kk;
      T fR = T(sc.fog1[0]) * kn + T(sc.fog2[0]) * kk;
      T fG = T(sc.fog1[1]) * kn + T(sc.fog2[1]) * kk;
      T fB = T(sc.fog1[2]) * kn + T(sc.fog2[2]) * kk;
      T k2 = distance / k.fog_c2_dist * kk;
      if (k2 > T(1)) k2 = 1;
      kn = T(1) - k2;
      fR = fR * kn + T(sc.fog3[0]) * k2;
      fG = fG * kn + T(sc.fog3[1]) * k2;
      fB = fB * kn + T(sc.fog3[2]) * k2;
      T fogDensity = T(0.3) * k.fog_density * densityTemp / (T(1) + k.fog_density * densityTemp);
      if (fogDensity > T(1)) fogDensity = 1;
      o.rgb.R = fogDensity * fR / T(65536) + (T(1) - fogDensity) * o.rgb.R;
      o.rgb.G = fogDensity * fG / T(65536) + (T(1) - fogDensity) * o.rgb.G;
      o.rgb.B = fogDensity * fB / T(65536) + (T(1) - fogDensity) * o.rgb.B;
      totalOpacity = (float)(fogDensity + (T(1) - fogDensity) * T(totalOpacity));
      o.alpha = fogDensity + (T(1) - fogDensity) * o.alpha;
    }

    if (vr) fx(1, index, point, dist_thresh, step, o, totalOpacity);

    if (totalOpacity > 1.0f) totalOpacity = 1.0f;
    if (o.alpha > T(1)) o.alpha = 1;
    o.opacity = T(totalOpacity);
  }
  return o;
}

// per-pixel view vector (src/Render3D.cpp:389-438); true when the fish-eye
// cut removes the pixel
__device__ __forceinline__ bool view_vector(const DevScene &sc, int x, int y, V3<double> &out) {
#pragma clang fp contract(off)
  const double tiles = sc.tiles;
  const double x2 = ((double)x / sc.width / tiles - 0.5 + sc.tile_x / tiles) * sc.aspect;
  const double z2 = ((double)y / sc.height / tiles - 0.5 + sc.tile_y / tiles);
  const double fov = sc.fov;
  bool cut = false;
  V3<double> v;
  if (sc.persp_type == 1) {
    if (sc.fish_eye_cut && sqrt(x2 * x2 + z2 * z2) > 0.5 / fov) cut = true;
    const double x3 = x2 * 3.14159265358979323846, z3 = z2 * 3.14159265358979323846;
    const double r = sqrt(x3 * x3 + z3 * z3);
    if (r == 0) v = v3<double>(0.0, 1.0, 0.0);
    else v = v3<double>(x3 / r * sin(r * fov), cos(r * fov), z3 / r * sin(r * fov));
    v = normalize(v);
  } else if (sc.persp_type == 2) {
    const double x3 = x2 * 3.14159265358979323846, z3 = z2 * 3.14159265358979323846;
    v = v3<double>(sin(fov * x3) * cos(fov * z3), cos(fov * x3) * cos(fov * z3), sin(fov * z3));
    v = normalize(v);
  } else {
    v = v3<double>(x2 * fov, 1.0, z2 * fov);
  }
  out = sc.R_identity ? v : mat_apply_nc(sc.R, v);
  return cut;
}

}  // namespace mbr
