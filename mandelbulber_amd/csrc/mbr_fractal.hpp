// mbr_fractal.hpp — distance estimator of the render core, templated on the
// arithmetic type (double: the FP64 path and the fixup bisection; float:
// the FP32 march and shading).
//
//   compute<T, MODE_NORMAL|MODE_COLOURING>  Compute<Mode>, src/fractal.cpp:40-1241
//       formulas trig_optim (:211-225), menger_sponge (:369-408),
//       tglad (:409-529); pre-folds (:125-181); termination (:1129-1191);
//       return values (:1193-1240)
//   calc_distance<T>                         CalculateDistance, :1245-1448
//                                            (analytic branch; primitives)
#pragma once

#include "mbr_core.hpp"
#include "mbr_scene_dev.hpp"

namespace mbr {

enum { MODE_NORMAL = 0, MODE_COLOURING = 1 };
enum { OBJ_FRACTAL = 0, OBJ_PLANE = 1, OBJ_WATER = 2, OBJ_SPHERE = 3, OBJ_SPHERE_INV = 4, OBJ_BOX = 5,
       OBJ_BOX_INV = 6 };

template <class T>
MBR_HD V3<T> ld3(const T *p) { return V3<T>{p[0], p[1], p[2]}; }

// z -> m*z with a full matrix product (rotinv matrices are never identity-
// flagged in the reference, src/algebra.cpp:202-217).
template <class T>
MBR_HD V3<T> rot9(const T *m, V3<T> v) { return mat_apply(m, v); }

template <class T, int Mode>
MBR_HD T compute(const FractalK<T> &f, V3<T> z, int N, int *iter_count) {
  using M = Math<T>;
  int L;
  T distance = 0;
  T r_dz = 1;
  const T p = f.power;
  T tgladDE = 1;
  T tgladColor = 1;
  const V3<T> constant = f.julia_mode ? ld3(f.julia) : z * f.constant_factor;
  T r = length(z);
  T min = Mode == MODE_COLOURING ? T(3.0e38) : T(0);  // min radius of the orbit (1e200 in FP64)
  if (Mode == MODE_COLOURING && sizeof(T) == 8) min = T(1e200);

  for (L = 0; L < N; L++) {
    if (f.tglad_fold) {
      const T lim = f.fold_limit, val = f.fold_value;
      if (z.x > lim) { z.x = val - z.x; tgladColor *= T(0.9); }
      else if (z.x < -lim) { z.x = -val - z.x; tgladColor *= T(0.9); }
      if (z.y > lim) { z.y = val - z.y; tgladColor *= T(0.9); }
      else if (z.y < -lim) { z.y = -val - z.y; tgladColor *= T(0.9); }
      if (z.z > lim) { z.z = val - z.z; tgladColor *= T(0.9); }
      else if (z.z < -lim) { z.z = -val - z.z; tgladColor *= T(0.9); }
      r = length(z);
    }
    if (f.sph_fold) {
      T r2 = r * r;
      if (r2 < f.sph_mR2) { z = z * f.sph_factor1; tgladDE *= f.sph_factor1; tgladColor += 1; }
      else if (r2 < f.sph_fR2) { T tf2 = f.sph_fR2 / r2; z = z * tf2; tgladDE *= tf2; tgladColor += 10; }
      r = length(z);
    }

    if (f.formula == 2) {  // trig_optim
      T th0 = M::asin(z.z / r);
      T ph0 = M::atan2(z.y, z.x);
      T rp = M::pow(r, p - T(1));
      T th = th0 * p;
      T ph = ph0 * p;
      T cth = M::cos(th);
      r_dz = rp * r_dz * p + T(1);
      rp *= r;
      z = V3<T>{cth * M::cos(ph), cth * M::sin(ph), M::sin(th)} * rp + constant;
      r = length(z);
    } else if (f.formula == 7) {  // menger_sponge
      T temp;
      z.x = M::fabs(z.x); z.y = M::fabs(z.y); z.z = M::fabs(z.z);
      if (z.x - z.y < 0) { temp = z.y; z.y = z.x; z.x = temp; }
      if (z.x - z.z < 0) { temp = z.z; z.z = z.x; z.x = temp; }
      if (z.y - z.z < 0) { temp = z.z; z.z = z.y; z.y = temp; }
      if (Mode == MODE_COLOURING) { T length2 = length(z); if (length2 < min) min = length2; }
      z = z * T(3);
      z.x -= T(2); z.y -= T(2);
      if (z.z > T(1)) z.z -= T(2);
      r = length(z);
      tgladDE *= T(3);
    } else if (f.formula == 8) {  // tglad (Mandelbox)
      const T lim = f.mb_limit, val = f.mb_value;
      if (f.mb_rot_on) {
        bool lockout = false;
        z = rot9(f.mb_rot[0][0], z);
        if (z.x > lim) { z.x = val - z.x; tgladColor += f.mb_col_X; lockout = true; }
        z = rot9(f.mb_rotinv[0][0], z);
        z = rot9(f.mb_rot[1][0], z);
        if (!lockout && z.x < -lim) { z.x = -val - z.x; tgladColor += f.mb_col_X; }
        z = rot9(f.mb_rotinv[1][0], z);
        lockout = false;
        z = rot9(f.mb_rot[0][1], z);
        if (z.y > lim) { z.y = val - z.y; tgladColor += f.mb_col_Y; lockout = true; }
        z = rot9(f.mb_rotinv[0][1], z);
        z = rot9(f.mb_rot[1][1], z);
        if (!lockout && z.y < -lim) { z.y = -val - z.y; tgladColor += f.mb_col_Y; }
        z = rot9(f.mb_rotinv[1][1], z);
        lockout = false;
        z = rot9(f.mb_rot[0][2], z);
        if (z.z > lim) { z.z = val - z.z; tgladColor += f.mb_col_Z; lockout = true; }
        z = rot9(f.mb_rotinv[0][2], z);
        z = rot9(f.mb_rot[1][2], z);
        if (!lockout && z.z < -lim) { z.z = -val - z.z; tgladColor += f.mb_col_Z; }
        z = rot9(f.mb_rotinv[1][2], z);
      } else {
        if (z.x > lim) { z.x = val - z.x; tgladColor += f.mb_col_X; }
        else if (z.x < -lim) { z.x = -val - z.x; tgladColor += f.mb_col_X; }
        if (z.y > lim) { z.y = val - z.y; tgladColor += f.mb_col_Y; }
        else if (z.y < -lim) { z.y = -val - z.y; tgladColor += f.mb_col_Y; }
        if (z.z > lim) { z.z = val - z.z; tgladColor += f.mb_col_Z; }
        else if (z.z < -lim) { z.z = -val - z.z; tgladColor += f.mb_col_Z; }
      }
      r = length(z);
      T r2 = r * r;
      z = z + ld3(f.mb_offset);
      if (r2 < f.mb_mR2) { z = z * f.mb_factor1; tgladDE *= f.mb_factor1; tgladColor += f.mb_col_Sp1; }
      else if (r2 < f.mb_fR2) { T tf2 = f.mb_fR2 / r2; z = z * tf2; tgladDE *= tf2; tgladColor += f.mb_col_Sp2; }
      z = z - ld3(f.mb_offset);
      if (f.mb_main_rot) z = mat_apply(f.mb_main, z);
      z = z * f.mb_scale + constant;
      tgladDE = tgladDE * f.mb_abs_scale + T(1);
      r = length(z);
    }

    // termination; "!(a <= b)" keeps an overflowed (NaN/inf) FP32 orbit
    // counted as escaped, and equals "a > b" for every finite value.
    if (f.formula == 7) {
      if (!(r <= T(1000))) { distance = (r - T(2)) / tgladDE; break; }
    } else if (f.formula == 8) {
      if (!(r <= T(1024))) { distance = r / M::fabs(tgladDE); break; }
    } else {
      if (Mode == MODE_NORMAL) {
        if (!(r <= f.bailout)) { distance = T(0.5) * r * M::log(r) / r_dz; break; }
      } else {
        distance = length(z);
        if (f.only_plane) {
          if (!(distance <= T(1e15))) {
            distance = (T(L) - M::log(M::log(r) / M::log(T(N))) / M::log(p)) / T(100);
            break;
          }
        } else {
          if (distance < min) min = distance;
          if (!(distance <= T(1e15))) { distance = min; break; }
        }
      }
    }
  }

  if (iter_count) *iter_count = L;
  if (Mode == MODE_NORMAL) {
    if (L == N) distance = 0;
    return distance;
  }
  if (f.formula == 8) return tgladColor * T(100) + length(z) * f.mb_col_R;
  if (f.formula == 7) return min * T(1000);
  return distance * T(5000);
}

// Primitives (src/primitives.cpp:15-119).  Box faces are axis planes, so
// the plane distances reduce exactly to coordinate differences.
template <class T>
MBR_HD T prim_box(V3<T> p, const T *c1, const T *c2) {
  T d = c1[0] - p.x;
  T t;
  t = p.x - c2[0]; d = (t > d) ? t : d;
  t = c1[1] - p.y; d = (t > d) ? t : d;
  t = p.y - c2[1]; d = (t > d) ? t : d;
  t = c1[2] - p.z; d = (t > d) ? t : d;
  t = p.z - c2[2]; d = (t > d) ? t : d;
  return d;
}
template <class T>
MBR_HD T prim_ibox(V3<T> p, const T *c1, const T *c2) {
  T d = p.x - c1[0];
  T t;
  t = c2[0] - p.x; d = (t < d) ? t : d;
  t = p.y - c1[1]; d = (t < d) ? t : d;
  t = c2[1] - p.y; d = (t < d) ? t : d;
  t = p.z - c1[2]; d = (t < d) ? t : d;
  t = c2[2] - p.z; d = (t < d) ? t : d;
  return d;
}
template <class T>
MBR_HD T prim_water(const FractalK<T> &f, V3<T> p) {
  using M = Math<T>;
  // plane (0,0,-1) through (0,0,height)
  T planeDistance = -(p.z - f.water_height);
  if (planeDistance < f.water_amplitude * T(10)) {
    if (f.water_rotated) p = mat_apply(f.water_rot, p);
    const T phase = f.water_phase;
    const T k = T(0.23);
    T waveX = 0, waveY = 0, pp = 1, p2 = T(0.05);
    for (int i = 1; i <= f.water_iterations; i++) {
      float p3 = (float)(pp * p2);
      T shift = phase / (T(i) / T(3) + T(1));
      T wx = M::sin(T(i) + T(0.4) * waveX * T(p3) + M::sin(k * p.y / f.water_length * T(p3)) +
                    p.x / f.water_length * T(p3) + shift) / pp;
      T wy = M::cos(T(i) + T(0.4) * waveY * T(p3) + M::sin(p.x / f.water_length * T(p3)) +
                    k * p.y / f.water_length * T(p3) + shift * T(0.23)) / pp;
      waveX += wx;
      waveY += wy;
      p2 = p2 + (T(1) - p2) * T(0.7);
      pp *= T(1.872);
    }
    planeDistance += (waveX + waveY) * f.water_amplitude;
  }
  return planeDistance;
}

struct DEInfo {
  int max_iter;
  int iters;
  int object;
  int work;  // formula iterations executed by this call
};

template <class T>
MBR_HD T calc_distance(const FractalK<T> &f, V3<T> point, T detailSize, int N, int minN, bool normalMode,
                       DEInfo &info) {
  info.object = OBJ_FRACTAL;
  info.max_iter = 0;
  info.work = 0;
  T distance;
  T limitBoxDist = 0;
  if (f.limits) {
    T da = point.x - f.limit_max[0], db = point.y - f.limit_max[1], dc = point.z - f.limit_max[2];
    T ta = -(point.x - f.limit_min[0]), tb = -(point.y - f.limit_min[1]), tc = -(point.z - f.limit_min[2]);
    da = (da > ta) ? da : ta;  // MAX()
    db = (db > tb) ? db : tb;
    dc = (dc > tc) ? dc : tc;
    limitBoxDist = (da > db) ? ((da > dc) ? da : dc) : ((db > dc) ? db : dc);  // dMax
    if (limitBoxDist > detailSize) return limitBoxDist;
  }
  if (!f.only_plane) {
    int L;
    distance = compute<T, MODE_NORMAL>(f, point, N, &L);
    info.work = (L < N) ? L + 1 : N;
    info.max_iter = (L == N);
    info.iters = L;
    if (distance < 0) distance = 0;
    if (L < minN && distance < detailSize) distance = detailSize;
    if (f.interior && !normalMode) {
      if (distance < T(0.5) * detailSize || L == N) { distance = detailSize; info.max_iter = 0; }
    } else if (f.interior && normalMode) {
      if (distance < T(0.9) * detailSize) { distance = detailSize - distance; info.max_iter = 0; }
    }
    if (f.iter_thresh && L < N) {
      if (distance < detailSize) distance = detailSize * T(1.01);
    }
  } else {
    distance = 10;
  }
  if (f.prim_plane) {
    T d = dot(ld3(f.plane_n), point - ld3(f.plane_c));
    if (!f.only_plane && d < distance) info.object = OBJ_PLANE;
    distance = (d < distance) ? d : distance;
  }
  if (f.prim_box) {
    T d = prim_box(point, f.box_c1, f.box_c2);
    if (d < distance) info.object = OBJ_BOX;
    distance = (d < distance) ? d : distance;
  }
  if (f.prim_ibox) {
    T d = prim_ibox(point, f.ibox_c1, f.ibox_c2);
    if (d < distance) info.object = OBJ_BOX_INV;
    distance = (d < distance) ? d : distance;
  }
  if (f.prim_sphere) {
    T d = length(point - ld3(f.sph_c)) - f.sph_r;
    if (d < distance) info.object = OBJ_SPHERE;
    distance = (d < distance) ? d : distance;
  }
  if (f.prim_isphere) {
    T d = f.isph_r - length(point - ld3(f.isph_c));
    if (d < distance) info.object = OBJ_SPHERE_INV;
    distance = (d < distance) ? d : distance;
  }
  if (f.prim_water) {
    T d = prim_water(f, point);
    if (d < distance) info.object = OBJ_WATER;
    distance = (d < distance) ? d : distance;
  }
  if (f.limits) {
    if (limitBoxDist < detailSize) distance = (distance > limitBoxDist) ? distance : limitBoxDist;
  }
  return distance;
}

}  // namespace mbr
