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

// the values of enumCalculationMode (src/fractal.h)
enum { MODE_NORMAL = 0, MODE_COLOURING = 1, MODE_DELTA1 = 3, MODE_DELTA2 = 4, MODE_ORBIT_TRAP = 5 };

// Formulas of the render core.  2, 7 and 8 have kernels of their own (F = 2,
// 7, 8); the others run through the generic FP64 path (F = 0): one
// CalculateDistance per DE call, analytic for trig_DE, by finite
// differences of the orbit radius (delta DE) for the rest.
MBR_HD bool formula_analytic(int formula) {
  return formula == 1 || formula == 2 || formula == 7 || formula == 8 || formula == 10 || formula == 17 ||
         formula == 18 || formula == 37;
}
// termination groups of src/fractal.cpp:1133-1149: Menger-like (r > 1000,
// (r - 2) / DE) and Mandelbox-like (r > 1024, r / |DE|)
MBR_HD bool formula_menger_group(int formula) { return formula == 7 || formula == 10; }
MBR_HD bool formula_box_group(int formula) {
  return formula == 8 || formula == 17 || formula == 18 || formula == 37;
}

// SmoothConditionAGreaterB / ALessB (src/common_math.h:87-88)
template <class T>
MBR_HD T smooth_gt(T a, T b, T sharp) { return T(1) / (T(1) + Math<T>::exp(sharp * (b - a))); }
template <class T>
MBR_HD T smooth_lt(T a, T b, T sharp) { return T(1) / (T(1) + Math<T>::exp(sharp * (a - b))); }
enum { OBJ_FRACTAL = 0, OBJ_PLANE = 1, OBJ_WATER = 2, OBJ_SPHERE = 3, OBJ_SPHERE_INV = 4, OBJ_BOX = 5,
       OBJ_BOX_INV = 6 };

template <class T>
MBR_HD V3<T> ld3(const T *p) { return V3<T>{p[0], p[1], p[2]}; }

// a.b and m*v with every product and sum rounded on its own (CVector3::Dot
// and CRotationMatrix::RotateVector as they compile on x86-64)
template <class T>
MBR_HD T dot_nc(V3<T> a, V3<T> b) {
#pragma clang fp contract(off)
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <class T>
MBR_HD V3<T> mat_apply_nc(const T *m, V3<T> v) {
#pragma clang fp contract(off)
  return V3<T>{m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z,
               m[6] * v.x + m[7] * v.y + m[8] * v.z};
}

// z -> m*z with a full matrix product (rotinv matrices are never identity-
// flagged in the reference, src/algebra.cpp:202-217).
template <class T>
MBR_HD V3<T> rot9(const T *m, V3<T> v) { return mat_apply(m, v); }

// Slot of the hybrid sequence at iteration L (CreateFormulaSequence,
// src/callbacks.cpp:1666-1703): slots 1-4 in turn for their iteration
// counts, then slot 5 for the rest (or, cyclic, for its count and the cycle
// repeats); -1 where the sequence holds no formula.
template <class T>
MBR_HD int hybrid_slot(const FractalK<T> &f, int L) {
  int pos = L;
  if (f.hyb_cyclic) {
    if (f.hyb_period <= 0) return -1;
    pos = L % f.hyb_period;
  }
  for (int k = 0; k < 4; k++) {
    if (f.hyb_f[k] == 0) continue;
    if (pos < f.hyb_it[k]) return k;
    pos -= f.hyb_it[k];
  }
  if (f.hyb_f[4] != 0 && (!f.hyb_cyclic || pos < f.hyb_it[4])) return 4;
  return -1;
}

// CVector3(alfa, beta) (src/algebra.hpp:33-38)
template <class T>
MBR_HD V3<T> angle_vector(T alfa, T beta) {
  using M = Math<T>;
  return V3<T>{M::cos(beta) * M::cos(alfa), M::cos(beta) * M::sin(alfa), M::sin(beta)};
}

// m = m * R_axis(angle) (CRotationMatrix::RotateX/Y/Z, src/algebra.cpp:94-143;
// m row-major), skipped for a zero angle as the reference does.
template <class T>
MBR_HD void rotate_right(T *m, bool &ident, int axis, T angle) {
#pragma clang fp contract(off)
  if (angle == T(0)) return;
  using M = Math<T>;
  const T s = M::sin(angle), c = M::cos(angle);
  T r[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (axis == 0) { r[0] = 1; r[4] = c; r[8] = c; r[5] = -s; r[7] = s; }
  else if (axis == 1) { r[4] = 1; r[8] = c; r[0] = c; r[6] = -s; r[2] = s; }
  else { r[8] = 1; r[0] = c; r[4] = c; r[1] = -s; r[3] = s; }
  T o[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) o[3 * i + j] = m[3 * i] * r[j] + m[3 * i + 1] * r[3 + j] + m[3 * i + 2] * r[6 + j];
  for (int k = 0; k < 9; k++) m[k] = o[k];
  ident = false;
}

// G: the formulas of the generic path are compiled in (always in FP64; in
// FP32 for the FP32 shading of the analytic generic formulas, whose kernels
// ask for them, so that the FP32 kernels of formulas 2, 7 and 8 stay small).
template <class T, int Mode, bool G = (sizeof(T) == 8)>
MBR_HD T compute(const FractalK<T> &f, V3<T> z, int N, int *iter_count) {
#pragma clang fp contract(off)
  using M = Math<T>;
  int L;
  T distance = 0;
  T r_dz = 1;
  V3<T> dz = V3<T>{T(1), T(0), T(0)};  // trig_DE's derivative vector (src/fractal.cpp:49-53)
  T ph_dz = 0, th_dz = 0;
  const int delta_limit = (Mode == MODE_DELTA2 && iter_count) ? *iter_count : -1;
  T w = 0, constantw = 0;  // fourth component of hypercomplex, quaternion, aexion
  T p = f.power;
  T mscale = f.mb_scale, mabs = f.mb_abs_scale;  // the hybrid sets scale = p
  T tgladDE = 1;
  T tgladColor = 1;
  T orbitTrapTotal = 0;  // the fake lights' orbit-trap sum
  V3<T> constant = f.julia_mode ? ld3(f.julia) : z * f.constant_factor;  // aexion rewrites it
  T r = length_nc(z);
  T min = Mode == MODE_COLOURING ? T(3.0e38) : T(0);  // min radius of the orbit (1e200 in FP64)
  if (Mode == MODE_COLOURING && sizeof(T) == 8) min = T(1e200);

  for (L = 0; L < N; L++) {
    int af = f.formula;  // the formula of this iteration (the hybrid sequence's, src/fractal.cpp:88-96)
    if constexpr (sizeof(T) == 8) {
      if (f.formula == 12) {
        // the sequence has N entries; longer runs (normals at 5N, colouring
        // at 10N) stay on its last entry (src/fractal.cpp:92-93)
        const int k = hybrid_slot(f, L < f.N ? L : f.N - 1);
        af = k >= 0 ? f.hyb_f[k] : 0;
        p = k >= 0 ? f.hyb_p[k] : T(0);
        mscale = p;
        mabs = M::fabs(p);
      }
    }
    if constexpr (G) {
      // IFS folding mode (src/fractal.cpp:101-127): the kaleidoscopic
      // vectors fold z without their intensities or edges; the scenes that
      // turn it on run the generic kernels
      if (f.ifs_fold) {
        if (f.ifs_abs[0]) z.x = M::fabs(z.x);
        if (f.ifs_abs[1]) z.y = M::fabs(z.y);
        if (f.ifs_abs[2]) z.z = M::fabs(z.z);
        for (int i = 0; i < 9; i++) {
          if (!f.ifs_en[i]) continue;
          if (f.ifs_rot_on[i]) z = mat_apply_nc(f.ifs_rot[i], z);
          const V3<T> dirv = ld3(f.ifs_dir[i]);
          const T len = dot_nc(z, dirv);
          if (len < f.ifs_dist[i]) z = z - dirv * (T(2) * (len - f.ifs_dist[i]));
        }
        const V3<T> off = ld3(f.ifs_offset);
        V3<T> zo = z - off;
        if (f.ifs_main_on) zo = mat_apply_nc(f.ifs_main, zo);
        z = zo + off;
        z = z * f.ifs_scale;
        z = z - off * (f.ifs_scale - T(1));
        r = length_nc(z);
      }
    }
    if (f.tglad_fold) {
      const T lim = f.fold_limit, val = f.fold_value;
      if (z.x > lim) { z.x = val - z.x; tgladColor *= T(0.9); }
      else if (z.x < -lim) { z.x = -val - z.x; tgladColor *= T(0.9); }
      if (z.y > lim) { z.y = val - z.y; tgladColor *= T(0.9); }
      else if (z.y < -lim) { z.y = -val - z.y; tgladColor *= T(0.9); }
      if (z.z > lim) { z.z = val - z.z; tgladColor *= T(0.9); }
      else if (z.z < -lim) { z.z = -val - z.z; tgladColor *= T(0.9); }
      r = length_nc(z);
    }
    if (f.sph_fold) {
      T r2 = r * r;
      if (r2 < f.sph_mR2) { z = z * f.sph_factor1; tgladDE *= f.sph_factor1; tgladColor += 1; }
      else if (r2 < f.sph_fR2) { T tf2 = f.sph_fR2 / r2; z = z * tf2; tgladDE *= tf2; tgladColor += 10; }
      r = length_nc(z);
    }

    if (af == 2) {  // trig_optim
      T th0 = M::asin(z.z / r);
      T ph0 = M::atan2(z.y, z.x);
      T rp = M::pow(r, p - T(1));
      T th = th0 * p;
      T ph = ph0 * p;
      T cth = M::cos(th);
      r_dz = rp * r_dz * p + T(1);
      rp *= r;
      z = V3<T>{cth * M::cos(ph), cth * M::sin(ph), M::sin(th)} * rp + constant;
      r = length_nc(z);
    } else if (af == 7) {  // menger_sponge
      T temp;
      z.x = M::fabs(z.x); z.y = M::fabs(z.y); z.z = M::fabs(z.z);
      if (z.x - z.y < 0) { temp = z.y; z.y = z.x; z.x = temp; }
      if (z.x - z.z < 0) { temp = z.z; z.z = z.x; z.x = temp; }
      if (z.y - z.z < 0) { temp = z.z; z.z = z.y; z.y = temp; }
      if (Mode == MODE_COLOURING) { T length2 = length_nc(z); if (length2 < min) min = length2; }
      z = z * T(3);
      z.x -= T(2); z.y -= T(2);
      if (z.z > T(1)) z.z -= T(2);
      r = length_nc(z);
      tgladDE *= T(3);
    } else if (af == 8) {  // tglad (Mandelbox)
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
      r = length_nc(z);
      T r2 = r * r;
      z = z + ld3(f.mb_offset);
      if (r2 < f.mb_mR2) { z = z * f.mb_factor1; tgladDE *= f.mb_factor1; tgladColor += f.mb_col_Sp1; }
      else if (r2 < f.mb_fR2) { T tf2 = f.mb_fR2 / r2; z = z * tf2; tgladDE *= tf2; tgladColor += f.mb_col_Sp2; }
      z = z - ld3(f.mb_offset);
      if (f.mb_main_rot) z = mat_apply(f.mb_main, z);
      z = z * mscale + constant;
      tgladDE = tgladDE * mabs + T(1);
      r = length_nc(z);
    } else if constexpr (!G) {
      // the generic formulas: in the kernels that ask for them only
    } else if (af == 1) {  // trig_DE (src/fractal.cpp:189-209)
      const T r1 = M::pow(r, p - T(1));
      const T r2 = r1 * r;
      const T th = M::atan2(z.y, z.x);                                   // GetAlpha
      const T ph = -M::atan2(z.z, M::sqrt(z.x * z.x + z.y * z.y));       // -GetBeta
      if (Mode == MODE_NORMAL) {
        const T p_r1_rdz = p * r1 * r_dz;
        const T ph_phdz = (p - T(1)) * ph + ph_dz;
        const T th_thdz = (p - T(1)) * th + th_dz;
        dz = angle_vector(th_thdz, ph_phdz) * p_r1_rdz + V3<T>{T(1), T(0), T(0)};
        r_dz = length_nc(dz);
        th_dz = M::atan2(dz.y, dz.x);
        ph_dz = -M::atan2(dz.z, M::sqrt(dz.x * dz.x + dz.y * dz.y));
      }
      z = angle_vector(p * th, p * ph) * r2 + constant;
      r = length_nc(z);
    } else if (af == 13) {  // mandelbulb2 (src/fractal.cpp:226-254)
      T temp, tempR;
      tempR = M::sqrt(z.x * z.x + z.y * z.y);
      z = z * (T(1) / tempR);
      temp = z.x * z.x - z.y * z.y;
      z.y = T(2) * z.x * z.y;
      z.x = temp;
      z = z * tempR;
      tempR = M::sqrt(z.y * z.y + z.z * z.z);
      z = z * (T(1) / tempR);
      temp = z.y * z.y - z.z * z.z;
      z.z = T(2) * z.y * z.z;
      z.y = temp;
      z = z * tempR;
      tempR = M::sqrt(z.x * z.x + z.z * z.z);
      z = z * (T(1) / tempR);
      temp = z.x * z.x - z.z * z.z;
      z.z = T(2) * z.x * z.z;
      z.x = temp;
      z = z * tempR;
      z = z * r;
      z = z + constant;
      r = length_nc(z);
    } else if (af == 14) {  // mandelbulb3 (src/fractal.cpp:255-282)
      T temp, tempR;
      T sign = 1, sign2 = 1;
      if (z.x < 0) sign2 = -1;
      tempR = M::sqrt(z.x * z.x + z.y * z.y);
      z = z * (T(1) / tempR);
      temp = z.x * z.x - z.y * z.y;
      z.y = T(2) * z.x * z.y;
      z.x = temp;
      z = z * tempR;
      if (z.x < 0) sign = -1;
      tempR = M::sqrt(z.x * z.x + z.z * z.z);
      z = z * (T(1) / tempR);
      temp = z.x * z.x - z.z * z.z;
      z.z = T(2) * z.x * z.z * sign2;
      z.x = temp * sign;
      z = z * tempR;
      z = z * r;
      z = z + constant;
      r = length_nc(z);
    } else if (af == 15) {  // mandelbulb4 (src/fractal.cpp:283-299)
      const T rp = M::pow(r, p - T(1));
      const T angZ = M::atan2(z.y, z.x);
      const T angY = M::atan2(z.z, z.x);
      const T angX = M::atan2(z.z, z.y);
      // CRotationMatrix: identity, then RotateX, RotateY, RotateZ, each a
      // right product; RotateVector skips a matrix no rotation touched
      T m[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
      bool ident = true;
      rotate_right<T>(m, ident, 0, angX * (p - T(1)));
      rotate_right<T>(m, ident, 1, angY * (p - T(1)));
      rotate_right<T>(m, ident, 2, angZ * (p - T(1)));
      if (!ident) z = V3<T>{m[0] * z.x + m[1] * z.y + m[2] * z.z, m[3] * z.x + m[4] * z.y + m[5] * z.z,
                            m[6] * z.x + m[7] * z.y + m[8] * z.z};
      z = z * rp + constant;
      r = length_nc(z);
    } else if (af == 11) {  // xenodreambuie (src/fractal.cpp:300-320)
      const T pi = T(3.14159265358979323846);
      const T rp = M::pow(r, p);
      const T th = M::atan2(z.y, z.x);
      T ph = M::acos(z.z / r);
      if (ph > T(0.5) * pi) ph = pi - ph;
      else if (ph < -T(0.5) * pi) ph = -pi - ph;
      z.x = rp * M::cos(th * p) * M::sin(ph * p);
      z.y = rp * M::sin(th * p) * M::sin(ph * p);
      z.z = rp * M::cos(ph * p);
      z = z + constant;
      r = length_nc(z);
    } else if (af == 3 || af == 6) {  // fast_trig, minus_fast_trig (src/fractal.cpp:321-350)
      const T x2 = z.x * z.x, y2 = z.y * z.y, z2 = z.z * z.z;
      const T temp = T(1) - z2 / (x2 + y2);
      const T newx = (x2 - y2) * temp;
      const T newy = T(2) * z.x * z.y * temp;
      const T newz = (af == 3 ? T(-2) : T(2)) * z.z * M::sqrt(x2 + y2);
      z.x = newx + constant.x;
      z.y = newy + constant.y;
      z.z = newz + constant.z;
      r = length_nc(z);
    } else if (af == 4) {  // hypercomplex (src/fractal.cpp:351-359)
      const V3<T> nz{z.x * z.x - z.y * z.y - z.z * z.z - w * w, T(2) * z.x * z.y - T(2) * w * z.z,
                     T(2) * z.x * z.z - T(2) * z.y * w};
      const T nw = T(2) * z.x * w - T(2) * z.y * z.z;
      z = nz + constant;
      w = nw;
      r = M::sqrt(z.x * z.x + z.y * z.y + z.z * z.z + w * w);
    } else if (af == 5) {  // quaternion (src/fractal.cpp:360-368)
      const V3<T> nz{z.x * z.x - z.y * z.y - z.z * z.z - w * w, T(2) * z.x * z.y, T(2) * z.x * z.z};
      const T nw = T(2) * z.x * w;
      z = nz + constant;
      w = nw;
      r = M::sqrt(z.x * z.x + z.y * z.y + z.z * z.z + w * w);
    } else if (af == 16) {  // foldingIntPow2 (src/fractal.cpp:794-836)
      const T ff = f.fip_factor;
      if (z.x > ff) z.x = ff * T(2) - z.x;
      else if (z.x < -ff) z.x = -ff * T(2) - z.x;
      if (z.y > ff) z.y = ff * T(2) - z.y;
      else if (z.y < -ff) z.y = -ff * T(2) - z.y;
      if (z.z > ff) z.z = ff * T(2) - z.z;
      else if (z.z < -ff) z.z = -ff * T(2) - z.z;
      r = length_nc(z);
      const T r2_2 = r * r;
      if (r2_2 < T(0.25)) z = z * (T(1) / T(0.25));
      else if (r2_2 < T(1)) z = z * (T(1) / r2_2);
      z = z * T(2);
      const T x2 = z.x * z.x, y2 = z.y * z.y, z2 = z.z * z.z;
      const T temp = T(1) - z2 / (x2 + y2);
      const T newx = (x2 - y2) * temp;
      const T newy = T(2) * z.x * z.y * temp;
      const T newz = T(-2) * z.z * M::sqrt(x2 + y2);
      z.x = newx + constant.x;
      z.y = newy + constant.y;
      z.z = newz + constant.z;
      z.z *= f.fip_zfactor;
      r = length_nc(z);
    } else if (af == 17) {  // smoothMandelbox (src/fractal.cpp:751-793)
      const T sm = f.mb_sharp, lim = f.mb_limit, val = f.mb_value;
      const T zk1 = smooth_gt(z.x, lim, sm), zk2 = smooth_lt(z.x, -lim, sm);
      z.x = z.x * (T(1) - zk1) + (val - z.x) * zk1;
      z.x = z.x * (T(1) - zk2) + (-val - z.x) * zk2;
      tgladColor += (zk1 + zk2) * f.mb_col_X;
      const T zk3 = smooth_gt(z.y, lim, sm), zk4 = smooth_lt(z.y, -lim, sm);
      z.y = z.y * (T(1) - zk3) + (val - z.y) * zk3;
      z.y = z.y * (T(1) - zk4) + (-val - z.y) * zk4;
      tgladColor += (zk3 + zk4) * f.mb_col_Y;
      const T zk5 = smooth_gt(z.z, lim, sm), zk6 = smooth_lt(z.z, -lim, sm);
      z.z = z.z * (T(1) - zk5) + (val - z.z) * zk5;
      z.z = z.z * (T(1) - zk6) + (-val - z.z) * zk6;
      tgladColor += (zk5 + zk6) * f.mb_col_Z;
      r = length_nc(z);
      const T r2 = r * r;
      const T tglad_factor2 = f.mb_fR2 / r2;
      const T rk1 = smooth_lt(r2, f.mb_mR2, sm);
      const T rk2 = smooth_lt(r2, f.mb_fR2, sm);
      const T rk21 = (T(1) - rk1) * rk2;
      z = z * (T(1) - rk1) + z * (f.mb_factor1 * rk1);
      z = z * (T(1) - rk21) + z * (tglad_factor2 * rk21);
      tgladDE = tgladDE * (T(1) - rk1) + tgladDE * (f.mb_factor1 * rk1);
      tgladDE = tgladDE * (T(1) - rk21) + tgladDE * (tglad_factor2 * rk21);
      tgladColor += rk1 * f.mb_col_Sp1;
      tgladColor += rk21 * f.mb_col_Sp2;
      if (f.mb_main_rot) z = mat_apply(f.mb_main, z);
      z = z * mscale + constant;
      tgladDE = tgladDE * mabs + T(1);
      r = length_nc(z);
    } else if (af == 19) {  // aexion (src/fractal.cpp:916-947)
      if (L == 0) {
        const T cx = M::fabs(constant.x + constant.y + constant.z) + f.c_add;
        const T cy = M::fabs(-constant.x - constant.y + constant.z) + f.c_add;
        const T cz = M::fabs(-constant.x + constant.y - constant.z) + f.c_add;
        const T cw = M::fabs(constant.x - constant.y - constant.z) + f.c_add;
        constant = V3<T>{cx, cy, cz};
        constantw = cw;
        const T tx = M::fabs(z.x + z.y + z.z) + f.c_add;
        const T ty = M::fabs(-z.x - z.y + z.z) + f.c_add;
        const T tz = M::fabs(-z.x + z.y - z.z) + f.c_add;
        const T tw = M::fabs(z.x - z.y - z.z) + f.c_add;
        z = V3<T>{tx, ty, tz};
        w = tw;
      }
      const T tx = z.x * z.x - z.y * z.y + T(2) * w * z.z + constant.x;
      const T ty = z.y * z.y - z.x * z.x + T(2) * w * z.z + constant.y;
      const T tz = z.z * z.z - w * w + T(2) * z.x * z.y + constant.z;
      const T tw = w * w - z.z * z.z + T(2) * z.x * z.y + constantw;
      z = V3<T>{tx, ty, tz};
      w = tw;
      r = M::sqrt(z.x * z.x + z.y * z.y + z.z * z.z + w * w);
    } else if (af == 20) {  // benesi (src/fractal.cpp:948-970)
      T r1 = z.y * z.y + z.z * z.z;
      T newx;
      if (constant.x < 0 || z.x < M::sqrt(r1)) newx = z.x * z.x - r1;
      else newx = -z.x * z.x + r1;
      r1 = -T(1) / M::sqrt(r1) * T(2) * M::fabs(z.x);
      const T newy = r1 * (z.y * z.y - z.z * z.z);
      const T newz = r1 * T(2) * z.y * z.z;
      z.x = newx + constant.x;
      z.y = newy + constant.y;
      z.z = newz + constant.z;
      r = length_nc(z);
    } else if (af == 10) {  // kaleidoscopic IFS (src/fractal.cpp:837-886)
      if (f.ifs_abs[0]) z.x = M::fabs(z.x);
      if (f.ifs_abs[1]) z.y = M::fabs(z.y);
      if (f.ifs_abs[2]) z.z = M::fabs(z.z);
      for (int i = 0; i < 9; i++) {
        if (!f.ifs_en[i]) continue;
        if (f.ifs_rot_on[i]) {
          const T *m = f.ifs_rot[i];
          z = V3<T>{m[0] * z.x + m[1] * z.y + m[2] * z.z, m[3] * z.x + m[4] * z.y + m[5] * z.z,
                    m[6] * z.x + m[7] * z.y + m[8] * z.z};
        }
        const V3<T> dirv = ld3(f.ifs_dir[i]);
        const T len = z.x * dirv.x + z.y * dirv.y + z.z * dirv.z;
        if (len < f.ifs_dist[i]) z = z - dirv * (T(2) * (len - f.ifs_dist[i]) * f.ifs_int[i]);
      }
      const V3<T> off = ld3(f.ifs_offset);
      V3<T> zo = z - off;
      if (f.ifs_main_on) {
        const T *m = f.ifs_main;
        zo = V3<T>{m[0] * zo.x + m[1] * zo.y + m[2] * zo.z, m[3] * zo.x + m[4] * zo.y + m[5] * zo.z,
                   m[6] * zo.x + m[7] * zo.y + m[8] * zo.z};
      }
      z = zo + off;
      if (f.ifs_edge[0] > 0) z.x = f.ifs_edge[0] - M::fabs(f.ifs_edge[0] - z.x);
      if (f.ifs_edge[1] > 0) z.y = f.ifs_edge[1] - M::fabs(f.ifs_edge[1] - z.y);
      if (f.ifs_edge[2] > 0) z.z = f.ifs_edge[2] - M::fabs(f.ifs_edge[2] - z.z);
      if (Mode == MODE_COLOURING) {
        const T length2 = length_nc(z);
        if (length2 < min) min = length2;
      }
      z = z * f.ifs_scale;
      const T sm1 = f.ifs_scale - T(1);
      if (f.ifs_menger) {
        z.x -= off.x * sm1;
        z.y -= off.y * sm1;
        if (z.z > T(0.5) * off.z * sm1) z.z -= off.z * sm1;
      } else {
        z = z - off * sm1;
      }
      tgladDE *= f.ifs_scale;
      r = length_nc(z);
    } else if (af >= 22 && af <= 36) {  // hybrid building blocks (src/fractal.cpp:984-1103)
      if (af == 22) z.x = z.x >= T(0) ? z.x * z.x / (M::fabs(z.x) + p) : -z.x * z.x / (M::fabs(z.x) + p);
      else if (af == 23) z.y = z.y >= T(0) ? z.y * z.y / (M::fabs(z.y) + p) : -z.y * z.y / (M::fabs(z.y) + p);
      else if (af == 24) z.z = z.z >= T(0) ? z.z * z.z / (M::fabs(z.z) + p) : -z.z * z.z / (M::fabs(z.z) + p);
      else if (af == 25) {  // invertR
        const T rInv = r * r / (r + p);
        z.x = z.x / r * rInv;
        z.y = z.y / r * rInv;
        z.z = z.z / r * rInv;
      } else if (af == 26) {  // sphericalFold (sets every component, as the reference)
        const T rr = r * r, pp = p * p;
        if (rr < pp) z = V3<T>{T(1) / pp, T(1) / pp, T(1) / pp};
        else if (rr < pp * T(4)) z = V3<T>{T(1) / rr, T(1) / rr, T(1) / rr};
      } else if (af == 27) {  // powXYZ
        z.x = z.x >= 0 ? M::pow(z.x, p) : -M::pow(-z.x, p);
        z.y = z.y >= 0 ? M::pow(z.y, p) : -M::pow(-z.y, p);
        z.z = z.z >= 0 ? M::pow(z.z, p) : -M::pow(-z.z, p);
      } else if (af == 28) z.x = z.x * p;
      else if (af == 29) z.y = z.y * p;
      else if (af == 30) z.z = z.z * p;
      else if (af == 31) z.x = z.x + p;
      else if (af == 32) z.y = z.y + p;
      else if (af == 33) z.z = z.z + p;
      else if (af == 34) {
        const T angle = M::atan2(z.z, z.y) * p;
        const T tempR = M::sqrt(z.z * z.z + z.y * z.y);
        z.y = tempR * M::cos(angle);
        z.z = tempR * M::sin(angle);
      } else if (af == 35) {
        const T angle = M::atan2(z.z, z.x) * p;
        const T tempR = M::sqrt(z.z * z.z + z.x * z.x);
        z.x = tempR * M::cos(angle);
        z.z = tempR * M::sin(angle);
      } else {
        const T angle = M::atan2(z.y, z.x) * p;
        const T tempR = M::sqrt(z.x * z.x + z.y * z.y);
        z.x = tempR * M::cos(angle);
        z.y = tempR * M::sin(angle);
      }
      r = length_nc(z);
    } else if (af == 21) {  // bristorbrot (src/fractal.cpp:971-983)
      const T newx = z.x * z.x - z.y * z.y - z.z * z.z;
      const T newy = z.y * (T(2) * z.x - z.z);
      const T newz = z.z * (T(2) * z.x + z.y);
      z.x = newx + constant.x;
      z.y = newy + constant.y;
      z.z = newz + constant.z;
      r = length_nc(z);
    } else if (af == 37) {  // generalizedFoldBox (src/fractal.cpp:531-749)
      // Walk the segment 0 -> z through the fold polyhedron X.Nv = solid:
      // the first face it crosses (mirror), the face the reflected segment
      // crosses next (rotation line), a third (inversion point); z folds
      // towards the last feature found unless it lies within the melt layer.
      const int sides = f.gf_sides;
      const T solid = f.gf_solid, melt = f.gf_melt;
      T a = 1;
      int side = -1;
      for (int i = 0; i < sides; i++) {
        const T c = dot_nc(z, ld3(f.gf_nv[i]));
        if (c > T(0) && a * c > solid) { side = i; a = solid / c; }
      }
      if (side != -1) {
        const int side_m = side;
        const V3<T> nm = ld3(f.gf_nv[side_m]);
        const V3<T> x_m = z - nm * (dot_nc(z, nm) - solid);
        V3<T> lv = x_m - nm;
        V3<T> y = nm;
        a = 1;
        side = -1;
        for (int i = 0; i < sides; i++) {
          if (i == side_m) continue;
          const V3<T> ni = ld3(f.gf_nv[i]);
          const T b = solid - dot_nc(y, ni);
          const T c = dot_nc(lv, ni);
          if (c > T(0) && a * c > b) { side = i; a = b / c; }
        }
        V3<T> target = x_m;
        T colour = f.mb_col_X;
        if (side != -1) {
          const V3<T> xi_mr = y + lv * a;
          const int side_r = side;
          const V3<T> nr = ld3(f.gf_nv[side_r]);
          // CVector3::Cross (src/algebra.hpp:93-100)
          V3<T> l_r{nm.y * nr.z - nm.z * nr.y, -nm.x * nr.z + nm.z * nr.x, nm.x * nr.y - nm.y * nr.x};
          T a_rmin = (dot_nc(z, l_r) - dot_nc(xi_mr, l_r)) / dot_nc(l_r, l_r);
          if (a_rmin < T(0)) { a_rmin = -a_rmin; l_r = l_r * T(-1); }
          const V3<T> x_r = xi_mr + l_r * a_rmin;
          lv = x_r - xi_mr;
          y = xi_mr;
          a = 1;
          side = -1;
          for (int i = 0; i < sides; i++) {
            if (i == side_m || i == side_r) continue;
            const V3<T> ni = ld3(f.gf_nv[i]);
            const T b = solid - dot_nc(y, ni);
            const T c = dot_nc(lv, ni);
            if (c > T(0) && a * c > b) { side = i; a = b / c; }
          }
          if (side != -1) {
            target = y + lv * a;
            colour = f.mb_col_Z;
          } else {
            target = x_r;
            colour = f.mb_col_Y;
          }
        }
        const V3<T> z2x = target - z;
        if (dot_nc(z2x, z2x) > melt * melt) {
          const T mag = length_nc(z2x);
          z = z + z2x * (T(2) * (mag - melt) / (mag + T(.00000001)));
          tgladColor += colour;
        }
      }
      r = length_nc(z);
      const T r2 = r * r;
      z = z + ld3(f.mb_offset);
      if (r2 < f.mb_mR2) { z = z * f.mb_factor1; tgladDE *= f.mb_factor1; tgladColor += f.mb_col_Sp1; }
      else if (r2 < f.mb_fR2) { T tf2 = f.mb_fR2 / r2; z = z * tf2; tgladDE *= tf2; tgladColor += f.mb_col_Sp2; }
      z = z - ld3(f.mb_offset);
      if (f.mb_main_rot) z = mat_apply(f.mb_main, z);
      z = z * mscale + constant;
      tgladDE = tgladDE * mabs + T(1);
      r = length_nc(z);
    } else if (af == 18) {  // mandelboxVaryScale4D (src/fractal.cpp:887-915)
      // the scale drifts each iteration; w is the fourth coordinate
      mscale = mscale + f.v4_scale_vary * (M::fabs(mscale) - T(1));
      const T fold = f.v4_fold, minr = f.v4_minr;
      const V3<T> old = z;
      z.x = M::fabs(z.x + fold) - M::fabs(z.x - fold) - z.x;
      z.y = M::fabs(z.y + fold) - M::fabs(z.y - fold) - z.y;
      z.z = M::fabs(z.z + fold) - M::fabs(z.z - fold) - z.z;
      w = M::fabs(w + fold) - M::fabs(w - fold) - w;
      if (z.x != old.x) tgladColor += f.mb_col_X;
      if (z.y != old.y) tgladColor += f.mb_col_Y;
      if (z.z != old.z) tgladColor += f.mb_col_Z;
      // pow(q, 1) is q in glibc's correctly rounded pow; the device pow can
      // be an ulp off, which the colouring orbit (10 N iterations) magnifies
      const T q = z.x * z.x + z.y * z.y + z.z * z.z + w * w;
      const T rr = f.v4_rpower == T(1) ? q : M::pow(q, f.v4_rpower);
      T m = mscale;
      if (rr < minr * minr) { m = mscale / (minr * minr); tgladColor += f.mb_col_Sp1; }
      else if (rr < T(1)) { m = mscale / rr; tgladColor += f.mb_col_Sp2; }
      z = z * m + constant;
      w = w * m + f.v4_wadd;
      tgladDE = tgladDE * M::fabs(m) + T(1);
      r = M::sqrt(z.x * z.x + z.y * z.y + z.z * z.z + w * w);
    }

    // delta-DE runs (src/fractal.cpp:1110-1119) stop on their own tests
    if (Mode == MODE_DELTA1) {
      if (r > T(1e10)) break;
    } else if (Mode == MODE_DELTA2) {
      if (L == delta_limit) break;
    }
    // the fake lights' orbit trap (src/fractal.cpp:1121-1130): inverse
    // squared distances to the trap point over iterations min..max, the sum
    // returned when the orbit leaves the trap's 1000 radius; otherwise the
    // last distance, or (as in every mode) the Menger or Mandelbox group's
    // escape value below
    if (Mode == MODE_ORBIT_TRAP) {
      distance = length_nc(z - ld3(f.fl_trap));
      if (L >= f.fl_min && L <= f.fl_max) orbitTrapTotal += T(1) / (distance * distance);
      if (distance > T(1000)) { distance = orbitTrapTotal; break; }
    }

    // termination; "!(a <= b)" keeps an overflowed (NaN/inf) FP32 orbit
    // counted as escaped, and equals "a > b" for every finite value.  The
    // formulas of the generic FP64 path keep the reference's "a > b" (a NaN
    // orbit never escapes).
    if (formula_menger_group(af)) {
      if (!(r <= T(1000))) { distance = (r - T(2)) / tgladDE; break; }
    } else if (formula_box_group(af)) {
      if (!(r <= T(1024))) { distance = r / M::fabs(tgladDE); break; }
    } else {
      if (Mode == MODE_NORMAL) {
        const bool esc = af == 2 ? !(r <= f.bailout) : r > f.bailout;
        if (esc) { distance = T(0.5) * r * M::log(r) / r_dz; break; }
      } else if (Mode == MODE_COLOURING) {
        distance = length_nc(z);
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
  if (Mode == MODE_DELTA1 || Mode == MODE_DELTA2) return r;
  if (Mode == MODE_ORBIT_TRAP) return distance;
  if (f.formula == 12) {  // hybrid colouring (src/fractal.cpp:1218-1229)
    if (min > T(100)) min = T(100);
    if (distance > T(20)) distance = T(20);
    if (tgladColor > T(1000)) tgladColor = T(1000);
    return distance * T(5000) + tgladColor * T(100) + min * T(1000);
  }
  if (formula_box_group(f.formula)) return tgladColor * T(100) + length_nc(z) * f.mb_col_R;
  if (formula_menger_group(f.formula)) return min * T(1000);
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

// ---------------------------------------------------------------------------
// Compute<normal> as a resumable state: de_begin sets up the orbit, each
// de_iter call runs one pass of the iteration loop body (pre-folds, formula,
// termination test) with the statements of src/fractal.cpp:90-1191 in order.
// Running de_iter to completion is Compute<normal>; the persistent kernels
// advance many such states one iteration at a time.
template <class T>
struct DEIter {
  V3<T> z, c;
  T r, aux;  // aux: r_dz (trig formulas) or tgladDE (folding formulas)
  T inv_r;   // 1/r, kept by the FP32 integer-power Mandelbulb path
  int L;
};

// IP: the Mandelbulb's integer power when the caller knows it at compile
// time (8, the standard bulb), else 0 (read from f.ipow).
template <class T, int F, bool P = false, bool A = (sizeof(T) < 8), int IP = 0>
MBR_HD void de_begin(const FractalK<T> &f, V3<T> p, DEIter<T> &s) {
  s.z = p;
  s.c = (!P && f.julia_mode) ? ld3(f.julia) : p * f.constant_factor;
  s.r = length(p);
  s.aux = 1;
  s.L = 0;
  s.inv_r = 0;
  if constexpr (F == 2 && A) {
    if (IP || f.ipow) {
      const T r2 = p.x * p.x + p.y * p.y + p.z * p.z;
      s.inv_r = fast_rsq(r2);
      s.r = r2 * s.inv_r;
    }
  }
}

// One loop pass.  Returns true when the call has finished; `dist` is then
// the return value of Compute<normal> (0 when the orbit stayed bounded for
// N iterations) and s.L the reference's iteration counter.
template <class T, int F, bool P = false, bool A = (sizeof(T) < 8), int IP = 0>
MBR_HD bool de_iter(const FractalK<T> &f, DEIter<T> &s, int N, T &dist) {
  using M = Math<T>;
  V3<T> z = s.z;
  T r = s.r;
  if ((!P && f.tglad_fold)) {
    const T lim = f.fold_limit, val = f.fold_value;
    if (z.x > lim) z.x = val - z.x;
    else if (z.x < -lim) z.x = -val - z.x;
    if (z.y > lim) z.y = val - z.y;
    else if (z.y < -lim) z.y = -val - z.y;
    if (z.z > lim) z.z = val - z.z;
    else if (z.z < -lim) z.z = -val - z.z;
    r = length(z);
  }
  if ((!P && f.sph_fold)) {
    T r2 = r * r;
    if (r2 < f.sph_mR2) { z = z * f.sph_factor1; s.aux *= f.sph_factor1; }
    else if (r2 < f.sph_fR2) { T tf2 = fdiv(f.sph_fR2, r2); z = z * tf2; s.aux *= tf2; }
    r = length(z);
  }
  bool esc;
  if constexpr (F == 2 && A) {
    if (IP || f.ipow) {
      // Integer power p (A: the FP32 kernels, and FP64 marching of the mixed
      // precision modes): the same map without inverse trig.  With
      // rho = |(x,y)|, (cos th, sin th) = (rho, z)/r and (cos ph, sin ph) =
      // (x, y)/rho, so cos/sin(p th) and cos/sin(p ph) are the parts of the
      // complex powers (rho + i z)^p / r^p and (x + i y)^p / rho^p
      // (de Moivre), formed by binary exponentiation.  1/r is carried from
      // the previous iteration; in FP32 reciprocals and square roots use the
      // hardware rsq/rcp.
      const int p = IP ? IP : f.ipow;
      T inv_r = s.inv_r;
      if ((!P && f.tglad_fold) || (!P && f.sph_fold)) inv_r = fast_rsq(z.x * z.x + z.y * z.y + z.z * z.z);
      const T rho2 = z.x * z.x + z.y * z.y;
      // e^{i th} = (rho + i z) / r and e^{i ph} = (x + i y) / rho; on the z
      // axis (rho = 0) cos th = 0 and atan2(+-0, x) is +-0 for x = +0 and
      // +-pi for x = -0 (selects, no branch)
      const bool off_axis = rho2 > T(0);
      const T inv_rho = fast_rsq(off_axis ? rho2 : T(1));
      T wre = off_axis ? z.x * inv_rho : Math<T>::copysign(T(1), z.x);
      T wim = off_axis ? z.y * inv_rho : T(0);
      if (p == 8) {
        // the standard Mandelbulb.  U = rho + i z is left unnormalised:
        // U^8 = r^8 e^{8 i th} carries the factor r^8 of the new point, so
        // z' = (Re U^8 cos 8ph, Re U^8 sin 8ph, Im U^8) + c.  U^2 is formed
        // from rho^2 directly; rho = 0 on the axis (inv_rho = 1 there).
        const T rho = rho2 * inv_rho;
        T ure = rho2 - z.z * z.z, uim = T(2) * rho * z.z;
        for (int k = 0; k < 2; k++) {
          const T t = ure * ure - uim * uim; uim = T(2) * ure * uim; ure = t;
        }
        for (int k = 0; k < 3; k++) {
          const T t = wre * wre - wim * wim; wim = T(2) * wre * wim; wre = t;
        }
        const T r2 = r * r, r4 = r2 * r2;
        s.aux = r4 * r4 * inv_r * s.aux * T(8) + T(1);  // 8 r^7 aux + 1
        z = V3<T>{ure * wre, ure * wim, uim} + s.c;
        const T zr2 = z.x * z.x + z.y * z.y + z.z * z.z;
        s.inv_r = fast_rsq(zr2);
        r = zr2 * s.inv_r;
        esc = !(r <= f.bailout);
        if (esc) dist = T(0.5) * r * fast_ln(r) * fast_rcp(s.aux);
        s.z = z;
        s.r = r;
        if (esc) return true;
        s.L++;
        if (s.L >= N) { dist = 0; return true; }
        return false;
      }
      T ure = off_axis ? rho2 * inv_rho * inv_r : T(0);
      T uim = z.z * inv_r;
      T pure = 1, puim = 0, pwre = 1, pwim = 0, rp1 = 1;
      {
        T rb = r;
        int e = p - 1;  // r^(p-1)
        while (e) {
          if (e & 1) rp1 *= rb;
          e >>= 1;
          if (e) rb *= rb;
        }
        e = p;
        while (e) {
          if (e & 1) {
            T t = pure * ure - puim * uim; puim = pure * uim + puim * ure; pure = t;
            t = pwre * wre - pwim * wim; pwim = pwre * wim + pwim * wre; pwre = t;
          }
          e >>= 1;
          if (e) {
            T t = ure * ure - uim * uim; uim = T(2) * ure * uim; ure = t;
            t = wre * wre - wim * wim; wim = T(2) * wre * wim; wre = t;
          }
        }
      }
      s.aux = rp1 * s.aux * T(p) + T(1);
      const T rp = rp1 * r;
      z = V3<T>{pure * pwre, pure * pwim, puim} * rp + s.c;
      const T r2 = z.x * z.x + z.y * z.y + z.z * z.z;
      s.inv_r = fast_rsq(r2);
      r = r2 * s.inv_r;
      esc = !(r <= f.bailout);
      if (esc) dist = T(0.5) * r * fast_ln(r) * fast_rcp(s.aux);
      s.z = z;
      s.r = r;
      if (esc) return true;
      s.L++;
      if (s.L >= N) { dist = 0; return true; }
      return false;
    }
  }
  if constexpr (F == 2) {  // trig_optim
    const T p = f.power;
    T th0 = M::asin(z.z / r);
    T ph0 = M::atan2(z.y, z.x);
    T rp = M::pow(r, p - T(1));
    T th = th0 * p;
    T ph = ph0 * p;
    T cth = M::cos(th);
    s.aux = rp * s.aux * p + T(1);
    rp *= r;
    z = V3<T>{cth * M::cos(ph), cth * M::sin(ph), M::sin(th)} * rp + s.c;
    r = length(z);
    esc = !(r <= f.bailout);
    if (esc) dist = T(0.5) * r * M::log(r) / s.aux;
  } else if constexpr (F == 7) {  // menger_sponge
    T temp;
    z.x = M::fabs(z.x); z.y = M::fabs(z.y); z.z = M::fabs(z.z);
    if (z.x - z.y < 0) { temp = z.y; z.y = z.x; z.x = temp; }
    if (z.x - z.z < 0) { temp = z.z; z.z = z.x; z.x = temp; }
    if (z.y - z.z < 0) { temp = z.z; z.z = z.y; z.y = temp; }
    z = z * T(3);
    z.x -= T(2); z.y -= T(2);
    if (z.z > T(1)) z.z -= T(2);
    r = length(z);
    s.aux *= T(3);
    esc = !(r <= T(1000));
    if (esc) dist = fdiv(r - T(2), s.aux);
  } else {  // tglad (Mandelbox)
    const T lim = f.mb_limit, val = f.mb_value;
    if (f.mb_rot_on) {
      bool lockout = false;
      z = rot9(f.mb_rot[0][0], z);
      if (z.x > lim) { z.x = val - z.x; lockout = true; }
      z = rot9(f.mb_rotinv[0][0], z);
      z = rot9(f.mb_rot[1][0], z);
      if (!lockout && z.x < -lim) z.x = -val - z.x;
      z = rot9(f.mb_rotinv[1][0], z);
      lockout = false;
      z = rot9(f.mb_rot[0][1], z);
      if (z.y > lim) { z.y = val - z.y; lockout = true; }
      z = rot9(f.mb_rotinv[0][1], z);
      z = rot9(f.mb_rot[1][1], z);
      if (!lockout && z.y < -lim) z.y = -val - z.y;
      z = rot9(f.mb_rotinv[1][1], z);
      lockout = false;
      z = rot9(f.mb_rot[0][2], z);
      if (z.z > lim) { z.z = val - z.z; lockout = true; }
      z = rot9(f.mb_rotinv[0][2], z);
      z = rot9(f.mb_rot[1][2], z);
      if (!lockout && z.z < -lim) z.z = -val - z.z;
      z = rot9(f.mb_rotinv[1][2], z);
    } else {
      if (z.x > lim) z.x = val - z.x;
      else if (z.x < -lim) z.x = -val - z.x;
      if (z.y > lim) z.y = val - z.y;
      else if (z.y < -lim) z.y = -val - z.y;
      if (z.z > lim) z.z = val - z.z;
      else if (z.z < -lim) z.z = -val - z.z;
    }
    r = length(z);
    T r2 = r * r;
    z = z + ld3(f.mb_offset);
    if (r2 < f.mb_mR2) { z = z * f.mb_factor1; s.aux *= f.mb_factor1; }
    else if (r2 < f.mb_fR2) { T tf2 = fdiv(f.mb_fR2, r2); z = z * tf2; s.aux *= tf2; }
    z = z - ld3(f.mb_offset);
    if (f.mb_main_rot) z = mat_apply(f.mb_main, z);
    z = z * f.mb_scale + s.c;
    s.aux = s.aux * f.mb_abs_scale + T(1);
    r = length(z);
    esc = !(r <= T(1024));
    if (esc) dist = fdiv(r, M::fabs(s.aux));
  }
  s.z = z;
  s.r = r;
  if (esc) return true;
  s.L++;
  if (s.L >= N) { dist = 0; return true; }
  return false;
}

// CalculateDistance (src/fractal.cpp:1245-1448), analytic branch, split into
// a start (limits early-out, primitive-only scenes) and a finish (clamp, minN
// rule, interior mode, iteration threshold, primitives, limits) around the
// resumable Compute<normal>.
template <class T>
struct DECall {
  DEIter<T> it;
  V3<T> p;
  T limitBox;
};

template <class T, bool P = false>
MBR_HD T de_finish(const FractalK<T> &f, V3<T> point, T distance, int L, bool fractal, int N, int minN,
                   bool normalMode, T detailSize, T limitBoxDist, DEInfo &info) {
  info.object = OBJ_FRACTAL;
  if (fractal) {
    info.work = (L < N) ? L + 1 : N;
    info.max_iter = (L == N);
    info.iters = L;
    if (distance < 0) distance = 0;
    if (L < minN && distance < detailSize) distance = detailSize;
    if ((!P && f.interior) && !normalMode) {
      if (distance < T(0.5) * detailSize || L == N) { distance = detailSize; info.max_iter = 0; }
    } else if ((!P && f.interior) && normalMode) {
      if (distance < T(0.9) * detailSize) { distance = detailSize - distance; info.max_iter = 0; }
    }
    if ((!P && f.iter_thresh) && L < N) {
      if (distance < detailSize) distance = detailSize * T(1.01);
    }
  } else {
    distance = 10;
    info.max_iter = 0;
    info.work = 0;
  }
  if ((!P && f.prim_plane)) {
    T d = dot(ld3(f.plane_n), point - ld3(f.plane_c));
    if (!(!P && f.only_plane) && d < distance) info.object = OBJ_PLANE;
    distance = (d < distance) ? d : distance;
  }
  if ((!P && f.prim_box)) {
    T d = prim_box(point, f.box_c1, f.box_c2);
    if (d < distance) info.object = OBJ_BOX;
    distance = (d < distance) ? d : distance;
  }
  if ((!P && f.prim_ibox)) {
    T d = prim_ibox(point, f.ibox_c1, f.ibox_c2);
    if (d < distance) info.object = OBJ_BOX_INV;
    distance = (d < distance) ? d : distance;
  }
  if ((!P && f.prim_sphere)) {
    T d = length(point - ld3(f.sph_c)) - f.sph_r;
    if (d < distance) info.object = OBJ_SPHERE;
    distance = (d < distance) ? d : distance;
  }
  if ((!P && f.prim_isphere)) {
    T d = f.isph_r - length(point - ld3(f.isph_c));
    if (d < distance) info.object = OBJ_SPHERE_INV;
    distance = (d < distance) ? d : distance;
  }
  if ((!P && f.prim_water)) {
    T d = prim_water(f, point);
    if (d < distance) info.object = OBJ_WATER;
    distance = (d < distance) ? d : distance;
  }
  if ((!P && f.limits)) {
    if (limitBoxDist < detailSize) distance = (distance > limitBoxDist) ? distance : limitBoxDist;
  }
  return distance;
}

// CalculateDistance (src/fractal.cpp:1245-1448) in one call for the formulas
// of the generic path: analytic (trig_DE) through Compute<normal>, the rest
// by the delta DE of :1307-1342 (four orbit radii, at the point and after the
// reference's in-place +-1e-10 updates along x, y, z).
// One orbit run of the generic path, kept out of line (one copy per mode).
template <class T, int Mode>
__attribute__((noinline)) __host__ __device__ T compute_generic(const FractalK<T> &f, V3<T> z, int N, int *iter_count) {
  return compute<T, Mode, true>(f, z, N, iter_count);
}

// Not inlined: the body (five orbit runs over every generic formula) would
// otherwise be copied into each of the shaders' distance-estimate sites.
template <class T, bool P = false>
__attribute__((noinline)) __host__ __device__ T calc_distance_generic(const FractalK<T> &f, V3<T> point, T detailSize, int N, int minN, bool normalMode,
                               DEInfo &info) {
#pragma clang fp contract(off)
  using M = Math<T>;
  info.object = OBJ_FRACTAL;
  info.max_iter = 0;
  info.work = 0;
  info.iters = 0;
  T limitBox = 0;
  if ((!P && f.limits)) {
    T da = point.x - f.limit_max[0], db = point.y - f.limit_max[1], dc = point.z - f.limit_max[2];
    T ta = -(point.x - f.limit_min[0]), tb = -(point.y - f.limit_min[1]), tc = -(point.z - f.limit_min[2]);
    da = (da > ta) ? da : ta;
    db = (db > tb) ? db : tb;
    dc = (dc > tc) ? dc : tc;
    limitBox = (da > db) ? ((da > dc) ? da : dc) : ((db > dc) ? db : dc);
    if (limitBox > detailSize) return limitBox;
  }
  if ((!P && f.only_plane))
    return de_finish<T, P>(f, point, T(10), 0, false, N, minN, normalMode, detailSize, limitBox, info);
  if (formula_analytic(f.formula)) {
    int L = 0;
    const T dist = compute_generic<T, MODE_NORMAL>(f, point, N, &L);
    return de_finish<T, P>(f, point, dist, L, true, N, minN, normalMode, detailSize, limitBox, info);
  }
  const T delta = T(1e-10);
  int L = 0;
  const T r = compute_generic<T, MODE_DELTA1>(f, point, N, &L);
  const int retval = L;
  V3<T> q = point;
  q.x += delta;
  T r2 = compute_generic<T, MODE_DELTA2>(f, q, N, &L);
  const T dr1 = M::fabs(r2 - r) / delta;
  q.x -= delta;
  q.y += delta;
  r2 = compute_generic<T, MODE_DELTA2>(f, q, N, &L);
  const T dr2 = M::fabs(r2 - r) / delta;
  q.y -= delta;
  q.z += delta;
  r2 = compute_generic<T, MODE_DELTA2>(f, q, N, &L);
  const T dr3 = M::fabs(r2 - r) / delta;
  const T dr = M::sqrt(dr1 * dr1 + dr2 * dr2 + dr3 * dr3);
  T dist = f.linear_de ? T(0.5) * r / dr : T(0.5) * r * M::log(r) / dr;
  if (dist < 0) dist = 0;
  if (retval == N) dist = 0;
  dist = de_finish<T, P>(f, point, dist, retval, true, N, minN, normalMode, detailSize, limitBox, info);
  info.work *= 4;  // four orbits
  return dist;
}

// Start one CalculateDistance; true when it finished without iterating.
template <class T, int F, bool P = false, bool A = (sizeof(T) < 8), int IP = 0>
MBR_HD bool de_call_begin(const FractalK<T> &f, V3<T> point, T detailSize, int N, int minN, bool normalMode,
                          DECall<T> &call, T &dist, DEInfo &info) {
  if constexpr (F == 0) {
    // the generic path runs the whole distance estimate in this call
    dist = calc_distance_generic<T, P>(f, point, detailSize, N, minN, normalMode, info);
    return true;
  }
  info.object = OBJ_FRACTAL;
  info.max_iter = 0;
  info.work = 0;
  info.iters = 0;
  call.p = point;
  call.limitBox = 0;
  if ((!P && f.limits)) {
    T da = point.x - f.limit_max[0], db = point.y - f.limit_max[1], dc = point.z - f.limit_max[2];
    T ta = -(point.x - f.limit_min[0]), tb = -(point.y - f.limit_min[1]), tc = -(point.z - f.limit_min[2]);
    da = (da > ta) ? da : ta;  // MAX()
    db = (db > tb) ? db : tb;
    dc = (dc > tc) ? dc : tc;
    call.limitBox = (da > db) ? ((da > dc) ? da : dc) : ((db > dc) ? db : dc);  // dMax
    if (call.limitBox > detailSize) { dist = call.limitBox; return true; }
  }
  if ((!P && f.only_plane)) {
    dist = de_finish<T, P>(f, point, T(10), 0, false, N, minN, normalMode, detailSize, call.limitBox, info);
    return true;
  }
  de_begin<T, F, P, A, IP>(f, point, call.it);
  if (N <= 0) {
    dist = de_finish<T, P>(f, point, T(0), 0, true, N, minN, normalMode, detailSize, call.limitBox, info);
    return true;
  }
  return false;
}

// Advance one iteration; true when the call finished (dist, info valid).
template <class T, int F, bool P = false, bool A = (sizeof(T) < 8), int IP = 0>
MBR_HD bool de_call_step(const FractalK<T> &f, DECall<T> &call, T detailSize, int N, int minN, bool normalMode,
                         T &dist, DEInfo &info) {
  T raw = 0;
  if (!de_iter<T, F, P, A, IP>(f, call.it, N, raw)) return false;
  dist = de_finish<T, P>(f, call.p, raw, call.it.L, true, N, minN, normalMode, detailSize, call.limitBox, info);
  return true;
}

template <class T, int F, bool A = (sizeof(T) < 8), int IP = 0>
MBR_HD T calc_distance_f(const FractalK<T> &f, V3<T> point, T detailSize, int N, int minN, bool normalMode,
                         DEInfo &info) {
  DECall<T> call;
  T dist = 0;
  if (de_call_begin<T, F, false, A, IP>(f, point, detailSize, N, minN, normalMode, call, dist, info)) return dist;
  while (!de_call_step<T, F, false, A, IP>(f, call, detailSize, N, minN, normalMode, dist, info)) {
  }
  return dist;
}

// Distance estimate at an FP64 point for an FP32 caller (camera-relative
// precision).  An FP32 iteration rounds z and c by about 6e-8 of their size;
// once the orbit's derivative aux has grown so that such a rounding moves the
// estimate by less than 1/170 of the threshold (aux * detail >= kMixedSwitch
// * max(|z|, |c|)), the rest of the iteration runs in FP32.  At ordinary
// zoom that holds at the start and the call is the FP32 one; at deep zoom
// (detail far below FP32 resolution of the coordinates) the first
// iterations run in FP64 from the exact point.
constexpr double kMixedSwitch = 1e-5;

template <int F>
MBR_HD bool de_call_begin_p64(const FractalK<double> &fd, const FractalK<float> &ff, V3<double> p, float detail,
                              int N, int minN, bool normalMode, DECall<float> &call, float &dist, DEInfo &info) {
  const V3<float> pf = cvt<float>(p);
  const float cmax = length(pf) * (ff.julia_mode ? 0.0f : (ff.constant_factor > 1.0f ? ff.constant_factor : 1.0f));
  const float zc = ff.julia_mode ? length(ld3(ff.julia)) : cmax;
  if ((double)detail >= kMixedSwitch * (double)(length(pf) > zc ? length(pf) : zc))
    return de_call_begin<float, F>(ff, pf, detail, N, minN, normalMode, call, dist, info);
  DECall<double> c64;
  double d64 = 0.0;
  if (de_call_begin<double, F>(fd, p, (double)detail, N, minN, normalMode, c64, d64, info)) {
    dist = (float)d64;
    return true;
  }
  const double cm = length(c64.it.c);
  for (;;) {
    const double zm = c64.it.r;
    if (Math<double>::fabs(c64.it.aux) * (double)detail >= kMixedSwitch * (zm > cm ? zm : cm)) break;
    if (de_call_step<double, F>(fd, c64, (double)detail, N, minN, normalMode, d64, info)) {
      dist = (float)d64;
      return true;
    }
  }
  call.p = pf;
  call.limitBox = (float)c64.limitBox;
  call.it.z = cvt<float>(c64.it.z);
  call.it.c = cvt<float>(c64.it.c);
  call.it.r = (float)c64.it.r;
  call.it.aux = (float)c64.it.aux;
  call.it.L = c64.it.L;
  const float r2 = call.it.z.x * call.it.z.x + call.it.z.y * call.it.z.y + call.it.z.z * call.it.z.z;
  call.it.inv_r = r2 > 0.0f ? fast_rsq(r2) : 0.0f;
  return false;
}

// Whole CalculateDistance at an FP64 point in the caller's precision T.  An
// FP32 caller whose detail is below FP32 resolution at the point runs the
// whole estimate in FP64: at deep zoom the switch to FP32 comes late in the
// orbit and at a different iteration in every lane, so a wave would pay for
// both paths (the Mandelbox surface kernel of a 1e-6-wide 4K frame took 1.45x
// the FP64 kernel's time that way).
template <class T, int F>
MBR_HD T calc_distance_p64(const FractalK<double> &fd, const FractalK<T> &f, V3<double> point, T detailSize, int N,
                           int minN, bool normalMode, DEInfo &info) {
  if constexpr (sizeof(T) == 8) {
    return calc_distance_f<double, F>(fd, point, detailSize, N, minN, normalMode, info);
  } else {
    const V3<float> pf = cvt<float>(point);
    const float cmax = length(pf) * (f.julia_mode ? 0.0f : (f.constant_factor > 1.0f ? f.constant_factor : 1.0f));
    const float zc = f.julia_mode ? length(ld3(f.julia)) : cmax;
    if ((double)detailSize < kMixedSwitch * (double)(length(pf) > zc ? length(pf) : zc))
      return (float)calc_distance_f<double, F>(fd, point, (double)detailSize, N, minN, normalMode, info);
    DECall<float> call;
    float dist = 0;
    if (de_call_begin_p64<F>(fd, f, point, detailSize, N, minN, normalMode, call, dist, info)) return dist;
    while (!de_call_step<float, F>(f, call, detailSize, N, minN, normalMode, dist, info)) {
    }
    return dist;
  }
}


template <class T>
MBR_HD T calc_distance(const FractalK<T> &f, V3<T> point, T detailSize, int N, int minN, bool normalMode,
                       DEInfo &info) {
  if (!f.ifs_fold) {  // the IFS folding pre-fold runs on the generic path only
    if (f.formula == 7) return calc_distance_f<T, 7>(f, point, detailSize, N, minN, normalMode, info);
    if (f.formula == 8) return calc_distance_f<T, 8>(f, point, detailSize, N, minN, normalMode, info);
    if (f.formula == 2) return calc_distance_f<T, 2>(f, point, detailSize, N, minN, normalMode, info);
  }
  return calc_distance_generic<T>(f, point, detailSize, N, minN, normalMode, info);
}

}  // namespace mbr
