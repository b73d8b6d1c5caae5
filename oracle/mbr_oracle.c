/*
 * mbr_oracle.c — CPU FP64 restatement of the Mandelbulber 1.21 render core.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP render core and the
 * timed CPU baseline.  It is never linked or called by the product.
 *
 * It restates, in plain C, the per-pixel path of the reference CPU renderer:
 *   Compute<normal|colouring>   src/fractal.cpp:40-1241 (formulas 2, 7, 8)
 *   CalculateDistance           src/fractal.cpp:1245-1448
 *   primitives                  src/primitives.cpp:15-119
 *   rotation matrices           src/algebra.cpp:94-171
 *   RayMarching                 src/Render3D.cpp:79-199
 *   MainThread per-pixel body   src/Render3D.cpp:236-333, :387-590
 *   shaders                     src/shaders.cpp:21-925, :1253-1305
 * in the same order of floating-point operations, with the reference's FP32
 * truncations (float dist_thresh, float totalOpacity, float framebuffer).
 *
 * Pinning (see tests/test_oracle_*.py): closed forms (far-field Mandelbulb
 * DE, Menger axis points, Mandelbox affine regime, exact primitive SDFs),
 * invariants (mirror and (p-1)-fold symmetry of the bulb, octahedral
 * symmetry of box and sponge), textbook reductions (power-2 bulb on z=0 is
 * the Mandelbrot set; integer-power bulb equals the complex-power form;
 * Menger membership by base-3 digits), brute force on tiny frames, and the
 * compiled reference (oracle/_ref: algebra.cpp, common_math.cpp,
 * primitives.cpp) for camera, projection and primitive conventions.
 */
#include "mbr_oracle.h"

#include <limits.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------ vec3 */
typedef struct { double x, y, z; } v3;

static inline v3 v3m(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 v3add(v3 a, v3 b) { return v3m(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 v3sub(v3 a, v3 b) { return v3m(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 v3mul(v3 a, double s) { return v3m(a.x * s, a.y * s, a.z * s); }
static inline double v3len(v3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
static inline double v3dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
/* CVector3::Normalize (src/algebra.hpp:104-111): multiply by 1/length */
static inline v3 v3norm(v3 a) { double n = 1.0 / v3len(a); return v3m(a.x * n, a.y * n, a.z * n); }
static inline v3 v3from(const double *p) { return v3m(p[0], p[1], p[2]); }
/* CVector3::Cross (src/algebra.hpp:93-100), its term order */
static inline v3 v3cross(v3 a, v3 b) { return v3m(a.y * b.z - a.z * b.y, -a.x * b.z + a.z * b.x, a.x * b.y - a.y * b.x); }

/* ------------------------------------------------------- rotation matrix */
/* CRotationMatrix (src/algebra.cpp:80-218): row-major m[0..8] = m11..m33,
 * rotations post-multiplied, each skipped when its angle is 0, and
 * RotateVector returns its argument unchanged while no rotation was applied. */
typedef struct { double m[9]; int zero; } rotm;

static void rot_null(rotm *r) {
  memset(r->m, 0, sizeof r->m);
  r->m[0] = r->m[4] = r->m[8] = 1.0;
  r->zero = 1;
}

static void mat_mul(double *a, const double *b) { /* a = a*b */
  double c[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      c[i * 3 + j] = a[i * 3 + 0] * b[0 * 3 + j] + a[i * 3 + 1] * b[1 * 3 + j] + a[i * 3 + 2] * b[2 * 3 + j];
  memcpy(a, c, sizeof c);
}

static void rot_x(rotm *r, double angle) {
  if (angle == 0.0) return;
  double s = sin(angle), c = cos(angle), t[9] = {0};
  t[0] = 1.0; t[4] = c; t[8] = c; t[5] = -s; t[7] = s;
  mat_mul(r->m, t); r->zero = 0;
}
static void rot_y(rotm *r, double angle) {
  if (angle == 0.0) return;
  double s = sin(angle), c = cos(angle), t[9] = {0};
  t[4] = 1.0; t[8] = c; t[0] = c; t[6] = -s; t[2] = s;
  mat_mul(r->m, t); r->zero = 0;
}
static void rot_z(rotm *r, double angle) {
  if (angle == 0.0) return;
  double s = sin(angle), c = cos(angle), t[9] = {0};
  t[8] = 1.0; t[0] = c; t[4] = c; t[1] = -s; t[3] = s;
  mat_mul(r->m, t); r->zero = 0;
}
static v3 rot_apply(const rotm *r, v3 v) {
  if (r->zero) return v;
  const double *m = r->m;
  return v3m(m[0] * v.x + m[1] * v.y + m[2] * v.z,
             m[3] * v.x + m[4] * v.y + m[5] * v.z,
             m[6] * v.x + m[7] * v.y + m[8] * v.z);
}
/* SetRotation(angles[3]) (src/algebra.cpp:145-151): Z(angles[2]) Y(angles[1]) X(angles[0]) */
static void rot_set3(rotm *r, const double a[3]) { rot_null(r); rot_z(r, a[2]); rot_y(r, a[1]); rot_x(r, a[0]); }
/* Transpose (src/algebra.cpp:202-217) marks the result as non-identity. */
static void rot_transpose(const rotm *r, rotm *out) {
  const double *m = r->m;
  double t[9] = {m[0], m[3], m[6], m[1], m[4], m[7], m[2], m[5], m[8]};
  memcpy(out->m, t, sizeof t); out->zero = 0;
}

/* ------------------------------------------------------------ derived set */
typedef struct {
  const mbr_scene *s;
  rotm mb_main, mb_rot[2][3], mb_rotinv[2][3];
  rotm ifs_main, ifs_rot[9];
  v3 ifs_dir[9];
  int *seq;        /* hybrid formula sequence, N entries (formulaSequence) */
  double *seq_pow; /* hybridPowerSequence */
  v3 fold_nv[20];  /* generalizedFoldBox: the fold polyhedron's face normals */
  int fold_sides;
} fractal_ctx;

/* The face normals of generalizedFoldBox's polyhedra, as RecalculateIFSParams
 * precomputes them (src/callbacks.cpp:1557-1664); type = genFoldBox.type. */
static int fold_normals(int type, v3 *nv) {
  const double t = 1.0 / sqrt(3.0);
  const v3 tet[4] = {{t, t, -t}, {t, -t, t}, {-t, t, t}, {-t, -t, -t}};
  const v3 oct4[4] = {{t, t, t}, {-t, -t, t}, {-t, t, -t}, {t, -t, -t}};
  const v3 cube[6] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
  int n = 0;
  switch (type) {
    case 1:  /* cube */
      for (int i = 0; i < 6; i++) nv[n++] = cube[i];
      return n;
    case 2:  /* octahedron */
    case 4:  /* octahedron + cube */
      for (int i = 0; i < 4; i++) nv[n++] = tet[i];
      for (int i = 0; i < 4; i++) nv[n++] = oct4[i];
      if (type == 4)
        for (int i = 0; i < 6; i++) nv[n++] = cube[i];
      return n;
    case 3: {  /* dodecahedron */
      const double aa = (1.0 + sqrt(5.0)) / 2.0, bb = 1.0 / sqrt(aa * aa + 1.0);
      const v3 d[12] = {{0, bb, aa * bb},  {0, bb, -aa * bb},  {0, -bb, aa * bb},  {0, -bb, -aa * bb},
                        {bb, aa * bb, 0},  {bb, -aa * bb, 0},  {-bb, aa * bb, 0},  {-bb, -aa * bb, 0},
                        {aa * bb, 0, bb},  {-aa * bb, 0, bb},  {aa * bb, 0, -bb},  {-aa * bb, 0, -bb}};
      for (int i = 0; i < 12; i++) nv[n++] = d[i];
      return n;
    }
    case 5: {  /* icosahedron */
      const double aa = (1.0 + sqrt(5.0)) / 2.0, ff = sqrt(aa * aa + 1.0 / (aa * aa));
      const double cc = aa / ff, dd = 1.0 / aa / ff;
      for (int i = 0; i < 4; i++) nv[n++] = tet[i];
      for (int i = 0; i < 4; i++) nv[n++] = oct4[i];
      const v3 ic[12] = {{0, dd, cc}, {0, dd, -cc}, {0, -dd, cc}, {0, -dd, -cc}, {cc, 0, dd}, {cc, 0, -dd},
                         {-cc, 0, dd}, {-cc, 0, -dd}, {dd, cc, 0}, {dd, -cc, 0}, {-dd, cc, 0}, {-dd, -cc, 0}};
      for (int i = 0; i < 12; i++) nv[n++] = ic[i];
      return n;
    }
    case 6:  /* hexagonal prism: the reference's angle step 3.14159 * 2 / 6 */
    case 7: {  /* pentagonal prism */
      const int k = type == 6 ? 6 : 5;
      const double th = 3.14159 * 2.0 / k;
      nv[n++] = v3m(0, 0, 1);
      nv[n++] = v3m(0, 0, -1);
      for (int i = 0; i < k; i++) nv[n++] = v3m(cos(th * i), sin(th * i), 0);
      return n;
    }
    default:  /* 0, tetrahedron */
      for (int i = 0; i < 4; i++) nv[n++] = tet[i];
      return n;
  }
}

static void fractal_ctx_init(fractal_ctx *f, const mbr_scene *s) {
  f->s = s;
  /* RecalculateIFSParams, IFS part (src/callbacks.cpp:1540-1546):
   * SetRotation(alfa, beta, gamma) = Z(alfa) Y(beta) X(gamma) (src/algebra.cpp:153-159) */
  rot_null(&f->ifs_main);
  rot_z(&f->ifs_main, s->ifs_rotation[0]); rot_y(&f->ifs_main, s->ifs_rotation[1]); rot_x(&f->ifs_main, s->ifs_rotation[2]);
  for (int i = 0; i < 9; i++) {
    rot_null(&f->ifs_rot[i]);
    rot_z(&f->ifs_rot[i], s->ifs_angles[i][0]); rot_y(&f->ifs_rot[i], s->ifs_angles[i][1]);
    rot_x(&f->ifs_rot[i], s->ifs_angles[i][2]);
    f->ifs_dir[i] = v3norm(v3from(s->ifs_direction[i]));
  }
  /* CreateFormulaSequence (src/callbacks.cpp:1666-1703), literally */
  f->seq = NULL;
  f->seq_pow = NULL;
  if (s->formula == MBR_FORMULA_HYBRID && s->N > 0) {
    const int N = s->N;
    f->seq = (int *)calloc((size_t)N, sizeof(int));
    f->seq_pow = (double *)calloc((size_t)N, sizeof(double));
    int number = 0;
    while (number < N) {
      for (int h = 0; h < 4; h++) {
        if (s->hybrid_formula[h] == 0) continue;
        for (int i = 0; i < s->hybrid_iterations[h]; i++) {
          if (number < N) { f->seq[number] = s->hybrid_formula[h]; f->seq_pow[number] = s->hybrid_power[h]; }
          number++;
        }
      }
      int temp_end = N;
      if (s->hybrid_cyclic) temp_end = s->hybrid_iterations[4];
      if (s->hybrid_formula[4] != 0)
        for (int i = 0; i < temp_end; i++) {
          if (number < N) { f->seq[number] = s->hybrid_formula[4]; f->seq_pow[number] = s->hybrid_power[4]; }
          number++;
        }
      if (!s->hybrid_cyclic) break;
    }
  }
  /* RecalculateIFSParams, Mandelbox part (src/callbacks.cpp:1548-1555) */
  f->fold_sides = fold_normals(s->mb_fold_mode, f->fold_nv);
  rot_set3(&f->mb_main, s->mb_rotation_main);
  for (int fold = 0; fold < 2; fold++)
    for (int axis = 0; axis < 3; axis++) {
      rot_set3(&f->mb_rot[fold][axis], s->mb_rotation[fold][axis]);
      rot_transpose(&f->mb_rot[fold][axis], &f->mb_rotinv[fold][axis]);
    }
}

/* ----------------------------------------------------------------- hash */
static inline uint32_t mix32(uint32_t x) {
  /* "lowbias32" integer finaliser (C. Wellons, hash-prospector) */
  x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
  return x;
}
uint32_t oracle_jitter_draw(uint32_t seed, uint32_t frame, uint32_t pixel, uint32_t step) {
  uint32_t h = mix32(seed ^ 0x9e3779b9U);
  h = mix32(h ^ frame);
  h = mix32(h ^ pixel);
  h = mix32(h ^ step);
  return h % 1001U;
}

/* double -> int as x86-64 cvttsd2si does it: out-of-range and NaN give
 * INT_MIN (the reference converts with a plain C cast). */
static inline int x86_d2i(double d) {
  if (!(d > -2147483649.0 && d < 2147483648.0)) return INT_MIN;
  return (int)d;
}
static inline int x86_abs(int v) { return v == INT_MIN ? INT_MIN : (v < 0 ? -v : v); }
/* double -> unsigned short via int, as gcc lowers (unsigned short)double */
static inline unsigned short x86_d2u16(double d) { return (unsigned short)(unsigned)x86_d2i(d); }
static inline unsigned char x86_d2u8(double d) { return (unsigned char)(unsigned)x86_d2i(d); }

/* --------------------------------------------------------------- Compute */
enum { MODE_NORMAL = 0, MODE_COLOURING = 1, MODE_DELTA1 = 3, MODE_DELTA2 = 4, MODE_ORBIT_TRAP = 5 };

/* CVector3(alfa, beta) (src/algebra.hpp:33-38) */
static v3 angle_vec(double alfa, double beta) { return v3m(cos(beta) * cos(alfa), cos(beta) * sin(alfa), sin(beta)); }

/* Compute<Mode> (src/fractal.cpp:40-1241) for trig_DE (:189-209),
 * trig_optim (:211-225), mandelbulb2/3/4 (:226-299), xenodreambuie
 * (:300-320), menger_sponge (:369-408) and tglad (:409-529) with the tglad
 * and spherical pre-folds (:125-181), the delta-DE stops (:1110-1119),
 * termination (:1129-1191) and return values (:1193-1240).  In
 * MODE_DELTA2, *iter_count is read as the iteration to stop at. */
static double compute(const fractal_ctx *f, int mode, v3 z, int N, int *iter_count) {
  const mbr_scene *s = f->s;
  int L;
  double distance = 0;
  double r_dz = 1;
  v3 dz = v3m(1.0, 0.0, 0.0);
  double ph_dz = 0, th_dz = 0;
  const int delta_limit = (mode == MODE_DELTA2 && iter_count) ? *iter_count : -1;
  double w = 0, constantw = 0;
  double p = s->power;
  double fixedRadius = s->mb_fixed_radius;
  double fR2 = fixedRadius * fixedRadius;
  double minRadius = s->mb_min_radius;
  double mR2 = minRadius * minRadius;
  double tglad_factor1 = fR2 / mR2;
  double tgladDE = 1.0;
  double orbitTrapTotal = 0.0;
  double scale = s->mb_scale;
  double tgladColor = 1.0;
  int formula = s->formula;
  v3 constant = s->julia_mode ? v3from(s->julia) : v3mul(z, s->constant_factor);
  double r = v3len(z);
  double min = 1e200;

  for (L = 0; L < N; L++) {
    if (f->seq) {  /* hybrid: this iteration's formula and power (src/fractal.cpp:88-96) */
      int tempL = L;
      if (tempL > s->N - 1) tempL = s->N - 1;
      formula = f->seq[tempL];
      p = f->seq_pow[tempL];
      scale = p;
    }
    if (s->ifs_folding_mode) {  /* src/fractal.cpp:100-126 */
      if (s->ifs_abs[0]) z.x = fabs(z.x);
      if (s->ifs_abs[1]) z.y = fabs(z.y);
      if (s->ifs_abs[2]) z.z = fabs(z.z);
      for (int i = 0; i < 9; i++) {
        if (!s->ifs_enabled[i]) continue;
        z = rot_apply(&f->ifs_rot[i], z);
        double length = v3dot(z, f->ifs_dir[i]);
        if (length < s->ifs_distance[i])
          z = v3sub(z, v3mul(f->ifs_dir[i], 2.0 * (length - s->ifs_distance[i])));
      }
      v3 off = v3from(s->ifs_offset);
      z = v3add(rot_apply(&f->ifs_main, v3sub(z, off)), off);
      z = v3mul(z, s->ifs_scale);
      z = v3sub(z, v3mul(off, s->ifs_scale - 1.0));
      r = v3len(z);
    }
    if (s->tglad_folding_mode) {
      double lim = s->folding_limit, val = s->folding_value;
      if (z.x > lim) { z.x = val - z.x; tgladColor *= 0.9; }
      else if (z.x < -lim) { z.x = -val - z.x; tgladColor *= 0.9; }
      if (z.y > lim) { z.y = val - z.y; tgladColor *= 0.9; }
      else if (z.y < -lim) { z.y = -val - z.y; tgladColor *= 0.9; }
      if (z.z > lim) { z.z = val - z.z; tgladColor *= 0.9; }
      else if (z.z < -lim) { z.z = -val - z.z; tgladColor *= 0.9; }
      r = v3len(z);
    }
    if (s->spherical_folding_mode) {
      double fR2_2 = s->spherical_folding_fixed * s->spherical_folding_fixed;
      double mR2_2 = s->spherical_folding_min * s->spherical_folding_min;
      double r2_2 = r * r;
      double tf1 = fR2_2 / mR2_2;
      if (r2_2 < mR2_2) { z = v3mul(z, tf1); tgladDE *= tf1; tgladColor += 1; }
      else if (r2_2 < fR2_2) { double tf2 = fR2_2 / r2_2; z = v3mul(z, tf2); tgladDE *= tf2; tgladColor += 10; }
      r = v3len(z);
    }

    switch (formula) {
      case MBR_FORMULA_TRIG_OPTIM: {
        double th0 = asin(z.z / r);
        double ph0 = atan2(z.y, z.x);
        double rp = pow(r, p - 1.0);
        double th = th0 * p;
        double ph = ph0 * p;
        double cth = cos(th);
        r_dz = rp * r_dz * p + 1.0;
        rp *= r;
        z = v3add(v3mul(v3m(cth * cos(ph), cth * sin(ph), sin(th)), rp), constant);
        r = v3len(z);
        break;
      }
      case MBR_FORMULA_MENGER_SPONGE: {
        double temp;
        z.x = fabs(z.x); z.y = fabs(z.y); z.z = fabs(z.z);
        if (z.x - z.y < 0) { temp = z.y; z.y = z.x; z.x = temp; }
        if (z.x - z.z < 0) { temp = z.z; z.z = z.x; z.x = temp; }
        if (z.y - z.z < 0) { temp = z.z; z.z = z.y; z.y = temp; }
        if (mode == MODE_COLOURING) { double length2 = v3len(z); if (length2 < min) min = length2; }
        z = v3mul(z, 3.0);
        z.x -= 2.0; z.y -= 2.0;
        if (z.z > 1.0) z.z -= 2.0;
        r = v3len(z);
        tgladDE *= 3.0;
        break;
      }
      case MBR_FORMULA_TGLAD: {
        double lim = s->mb_folding_limit, val = s->mb_folding_value;
        if (s->mb_rotations_enabled) {
          int lockout = 0;
          z = rot_apply(&f->mb_rot[0][0], z);
          if (z.x > lim) { z.x = val - z.x; tgladColor += s->mb_color_X; lockout = 1; }
          z = rot_apply(&f->mb_rotinv[0][0], z);
          z = rot_apply(&f->mb_rot[1][0], z);
          if (!lockout && z.x < -lim) { z.x = -val - z.x; tgladColor += s->mb_color_X; }
          z = rot_apply(&f->mb_rotinv[1][0], z);
          lockout = 0;
          z = rot_apply(&f->mb_rot[0][1], z);
          if (z.y > lim) { z.y = val - z.y; tgladColor += s->mb_color_Y; lockout = 1; }
          z = rot_apply(&f->mb_rotinv[0][1], z);
          z = rot_apply(&f->mb_rot[1][1], z);
          if (!lockout && z.y < -lim) { z.y = -val - z.y; tgladColor += s->mb_color_Y; }
          z = rot_apply(&f->mb_rotinv[1][1], z);
          lockout = 0;
          z = rot_apply(&f->mb_rot[0][2], z);
          if (z.z > lim) { z.z = val - z.z; tgladColor += s->mb_color_Z; lockout = 1; }
          z = rot_apply(&f->mb_rotinv[0][2], z);
          z = rot_apply(&f->mb_rot[1][2], z);
          if (!lockout && z.z < -lim) { z.z = -val - z.z; tgladColor += s->mb_color_Z; }
          z = rot_apply(&f->mb_rotinv[1][2], z);
        } else {
          if (z.x > lim) { z.x = val - z.x; tgladColor += s->mb_color_X; }
          else if (z.x < -lim) { z.x = -val - z.x; tgladColor += s->mb_color_X; }
          if (z.y > lim) { z.y = val - z.y; tgladColor += s->mb_color_Y; }
          else if (z.y < -lim) { z.y = -val - z.y; tgladColor += s->mb_color_Y; }
          if (z.z > lim) { z.z = val - z.z; tgladColor += s->mb_color_Z; }
          else if (z.z < -lim) { z.z = -val - z.z; tgladColor += s->mb_color_Z; }
        }
        r = v3len(z);
        double r2 = r * r;
        z = v3add(z, v3from(s->mb_offset));
        if (r2 < mR2) { z = v3mul(z, tglad_factor1); tgladDE *= tglad_factor1; tgladColor += s->mb_color_Sp1; }
        else if (r2 < fR2) { double tf2 = fR2 / r2; z = v3mul(z, tf2); tgladDE *= tf2; tgladColor += s->mb_color_Sp2; }
        z = v3sub(z, v3from(s->mb_offset));
        z = rot_apply(&f->mb_main, z);
        z = v3add(v3mul(z, scale), constant);
        tgladDE = tgladDE * fabs(scale) + 1.0;
        r = v3len(z);
        break;
      }
      case MBR_FORMULA_TRIG_DE: {
        double r1 = pow(r, p - 1);
        double r2 = r1 * r;
        double th = atan2(z.y, z.x);
        double ph = -atan2(z.z, sqrt(z.x * z.x + z.y * z.y));
        if (mode == MODE_NORMAL) {
          double p_r1_rdz = p * r1 * r_dz;
          double ph_phdz = (p - 1.0) * ph + ph_dz;
          double th_thdz = (p - 1.0) * th + th_dz;
          dz = v3add(v3mul(angle_vec(th_thdz, ph_phdz), p_r1_rdz), v3m(1.0, 0.0, 0.0));
          r_dz = v3len(dz);
          th_dz = atan2(dz.y, dz.x);
          ph_dz = -atan2(dz.z, sqrt(dz.x * dz.x + dz.y * dz.y));
        }
        z = v3add(v3mul(angle_vec(p * th, p * ph), r2), constant);
        r = v3len(z);
        break;
      }
      case MBR_FORMULA_MANDELBULB2: {
        double temp, tempR;
        tempR = sqrt(z.x * z.x + z.y * z.y);
        z = v3mul(z, 1.0 / tempR);
        temp = z.x * z.x - z.y * z.y; z.y = 2.0 * z.x * z.y; z.x = temp;
        z = v3mul(z, tempR);
        tempR = sqrt(z.y * z.y + z.z * z.z);
        z = v3mul(z, 1.0 / tempR);
        temp = z.y * z.y - z.z * z.z; z.z = 2.0 * z.y * z.z; z.y = temp;
        z = v3mul(z, tempR);
        tempR = sqrt(z.x * z.x + z.z * z.z);
        z = v3mul(z, 1.0 / tempR);
        temp = z.x * z.x - z.z * z.z; z.z = 2.0 * z.x * z.z; z.x = temp;
        z = v3mul(z, tempR);
        z = v3mul(z, r);
        z = v3add(z, constant);
        r = v3len(z);
        break;
      }
      case MBR_FORMULA_MANDELBULB3: {
        double temp, tempR, sign = 1.0, sign2 = 1.0;
        if (z.x < 0) sign2 = -1.0;
        tempR = sqrt(z.x * z.x + z.y * z.y);
        z = v3mul(z, 1.0 / tempR);
        temp = z.x * z.x - z.y * z.y; z.y = 2.0 * z.x * z.y; z.x = temp;
        z = v3mul(z, tempR);
        if (z.x < 0) sign = -1.0;
        tempR = sqrt(z.x * z.x + z.z * z.z);
        z = v3mul(z, 1.0 / tempR);
        temp = z.x * z.x - z.z * z.z; z.z = 2.0 * z.x * z.z * sign2; z.x = temp * sign;
        z = v3mul(z, tempR);
        z = v3mul(z, r);
        z = v3add(z, constant);
        r = v3len(z);
        break;
      }
      case MBR_FORMULA_MANDELBULB4: {
        double rp = pow(r, p - 1);
        double angZ = atan2(z.y, z.x), angY = atan2(z.z, z.x), angX = atan2(z.z, z.y);
        rotm rm; rot_null(&rm);
        rot_x(&rm, angX * (p - 1));
        rot_y(&rm, angY * (p - 1));
        rot_z(&rm, angZ * (p - 1));
        z = v3add(v3mul(rot_apply(&rm, z), rp), constant);
        r = v3len(z);
        break;
      }
      case MBR_FORMULA_FAST_TRIG:
      case MBR_FORMULA_MINUS_FAST_TRIG: {  /* src/fractal.cpp:321-350 */
        double x2 = z.x * z.x, y2 = z.y * z.y, z2 = z.z * z.z;
        double temp = 1.0 - z2 / (x2 + y2);
        double newx = (x2 - y2) * temp;
        double newy = 2.0 * z.x * z.y * temp;
        double newz = (formula == MBR_FORMULA_FAST_TRIG ? -2.0 : 2.0) * z.z * sqrt(x2 + y2);
        z.x = newx + constant.x; z.y = newy + constant.y; z.z = newz + constant.z;
        r = v3len(z);
        break;
      }
      case MBR_FORMULA_HYPERCOMPLEX: {  /* :351-359 */
        v3 nz = v3m(z.x * z.x - z.y * z.y - z.z * z.z - w * w, 2.0 * z.x * z.y - 2.0 * w * z.z,
                    2.0 * z.x * z.z - 2.0 * z.y * w);
        double nw = 2.0 * z.x * w - 2.0 * z.y * z.z;
        z = v3add(nz, constant);
        w = nw;
        r = sqrt(z.x * z.x + z.y * z.y + z.z * z.z + w * w);
        break;
      }
      case MBR_FORMULA_QUATERNION: {  /* :360-368 */
        v3 nz = v3m(z.x * z.x - z.y * z.y - z.z * z.z - w * w, 2.0 * z.x * z.y, 2.0 * z.x * z.z);
        double nw = 2.0 * z.x * w;
        z = v3add(nz, constant);
        w = nw;
        r = sqrt(z.x * z.x + z.y * z.y + z.z * z.z + w * w);
        break;
      }
      case MBR_FORMULA_FOLDING_INT_POW2: {  /* :794-836 */
        double ff = s->folding_int_pow_factor;
        if (z.x > ff) z.x = ff * 2.0 - z.x; else if (z.x < -ff) z.x = -ff * 2.0 - z.x;
        if (z.y > ff) z.y = ff * 2.0 - z.y; else if (z.y < -ff) z.y = -ff * 2.0 - z.y;
        if (z.z > ff) z.z = ff * 2.0 - z.z; else if (z.z < -ff) z.z = -ff * 2.0 - z.z;
        r = v3len(z);
        double fR2_2 = 1.0, mR2_2 = 0.25, r2_2 = r * r, tf1_2 = fR2_2 / mR2_2;
        if (r2_2 < mR2_2) z = v3mul(z, tf1_2);
        else if (r2_2 < fR2_2) { double tf2_2 = fR2_2 / r2_2; z = v3mul(z, tf2_2); }
        z = v3mul(z, 2.0);
        double x2 = z.x * z.x, y2 = z.y * z.y, z2 = z.z * z.z;
        double temp = 1.0 - z2 / (x2 + y2);
        double newx = (x2 - y2) * temp;
        double newy = 2.0 * z.x * z.y * temp;
        double newz = -2.0 * z.z * sqrt(x2 + y2);
        z.x = newx + constant.x; z.y = newy + constant.y; z.z = newz + constant.z;
        z.z *= s->folding_int_pow_z_factor;
        r = v3len(z);
        break;
      }
      case MBR_FORMULA_GENERALIZED_FOLD_BOX: {  /* :531-749 */
        const v3 *Nv = f->fold_nv;
        const int sides = f->fold_sides;
        const double melt = s->mb_melt, solid = s->mb_solid;
        /* the closest fold plane X.Nv = solid cutting the segment origin..z */
        v3 Lv = z, Y = v3m(0, 0, 0);
        double a = 1, b, c;
        int side = -1;
        for (int i = 0; i < sides; i++) {
          b = solid;
          c = v3dot(Lv, Nv[i]);
          if ((c > 0) && ((a * c) > b)) { side = i; a = b / c; }
        }
        if (side != -1) {  /* mirror check */
          const int side_m = side;
          const v3 Nv_m = Nv[side_m];
          const v3 X_m = v3sub(z, v3mul(Nv_m, v3dot(z, Nv_m) - solid));
          Lv = v3sub(X_m, Nv_m);
          Y = Nv_m;
          a = 1;
          side = -1;
          for (int i = 0; i < sides; i++) {
            if (i != side_m) {
              b = solid - v3dot(Y, Nv[i]);
              c = v3dot(Lv, Nv[i]);
              if ((c > 0) && ((a * c) > b)) { side = i; a = b / c; }
            }
          }
          if (side != -1) {  /* rotation check */
            const v3 Xmr_intersect = v3add(Y, v3mul(Lv, a));
            const int side_r = side;
            const v3 Nv_r = Nv[side_r];
            v3 L_r = v3cross(Nv_m, Nv_r);
            double a_rmin = (v3dot(z, L_r) - v3dot(Xmr_intersect, L_r)) / v3dot(L_r, L_r);
            if (a_rmin < 0) { a_rmin = -a_rmin; L_r = v3mul(L_r, -1.0); }
            const v3 X_r = v3add(Xmr_intersect, v3mul(L_r, a_rmin));
            Lv = v3sub(X_r, Xmr_intersect);
            Y = Xmr_intersect;
            a = 1;
            side = -1;
            for (int i = 0; i < sides; i++) {
              if ((i != side_m) && (i != side_r)) {
                b = solid - v3dot(Y, Nv[i]);
                c = v3dot(Lv, Nv[i]);
                if ((c > 0) && ((a * c) > b)) { side = i; a = b / c; }
              }
            }
            if (side != -1) {  /* inversion point */
              const v3 X_i = v3add(Y, v3mul(Lv, a));
              const v3 z2X = v3sub(X_i, z);
              if (v3dot(z2X, z2X) > (melt * melt)) {
                double z2X_mag = v3len(z2X);
                z = v3add(z, v3mul(z2X, 2 * (z2X_mag - melt) / (z2X_mag + .00000001)));
                tgladColor += s->mb_color_Z;
              }
            } else {  /* rotation line */
              const v3 z2X = v3sub(X_r, z);
              if (v3dot(z2X, z2X) > (melt * melt)) {
                double z2X_mag = v3len(z2X);
                z = v3add(z, v3mul(z2X, 2 * (z2X_mag - melt) / (z2X_mag + .00000001)));
                tgladColor += s->mb_color_Y;
              }
            }
          } else {  /* mirror plane */
            const v3 z2X = v3sub(X_m, z);
            if (v3dot(z2X, z2X) > (melt * melt)) {
              double z2X_mag = v3len(z2X);
              z = v3add(z, v3mul(z2X, 2 * (z2X_mag - melt) / (z2X_mag + .00000001)));
              tgladColor += s->mb_color_X;
            }
          }
        }
        r = v3len(z);
        double r2 = r * r;
        z = v3add(z, v3from(s->mb_offset));
        if (r2 < mR2) { z = v3mul(z, tglad_factor1); tgladDE *= tglad_factor1; tgladColor += s->mb_color_Sp1; }
        else if (r2 < fR2) { double tf2 = fR2 / r2; z = v3mul(z, tf2); tgladDE *= tf2; tgladColor += s->mb_color_Sp2; }
        z = v3sub(z, v3from(s->mb_offset));
        z = rot_apply(&f->mb_main, z);
        z = v3add(v3mul(z, scale), constant);
        tgladDE = tgladDE * fabs(scale) + 1.0;
        r = v3len(z);
        break;
      }
      case MBR_FORMULA_MANDELBOX_VARY_SCALE_4D: {  /* :887-915 */
        const double fold = s->mb_vary_fold, minR = s->mb_vary_minr;
        scale = scale + s->mb_vary_scale_vary * (fabs(scale) - 1.0);
        v3 oldz = z;
        z.x = fabs(z.x + fold) - fabs(z.x - fold) - z.x;
        z.y = fabs(z.y + fold) - fabs(z.y - fold) - z.y;
        z.z = fabs(z.z + fold) - fabs(z.z - fold) - z.z;
        w = fabs(w + fold) - fabs(w - fold) - w;
        if (z.x != oldz.x) tgladColor += s->mb_color_X;
        if (z.y != oldz.y) tgladColor += s->mb_color_Y;
        if (z.z != oldz.z) tgladColor += s->mb_color_Z;
        double rr = pow(z.x * z.x + z.y * z.y + z.z * z.z + w * w, s->mb_vary_rpower);
        double m = scale;
        if (rr < minR * minR) { m = scale / (minR * minR); tgladColor += s->mb_color_Sp1; }
        else if (rr < 1.0) { m = scale / rr; tgladColor += s->mb_color_Sp2; }
        z = v3add(v3mul(z, m), constant);
        w = w * m + s->mb_vary_wadd;
        tgladDE = tgladDE * fabs(m) + 1.0;
        r = sqrt(z.x * z.x + z.y * z.y + z.z * z.z + w * w);
        break;
      }
      case MBR_FORMULA_SMOOTH_MANDELBOX: {  /* :751-793 */
        double sm = s->mb_sharpness, lim = s->mb_folding_limit, val = s->mb_folding_value;
#define SGT(a, b) (1.0 / (1.0 + exp(sm * ((b) - (a)))))
#define SLT(a, b) (1.0 / (1.0 + exp(sm * ((a) - (b)))))
        double zk1 = SGT(z.x, lim), zk2 = SLT(z.x, -lim);
        z.x = z.x * (1.0 - zk1) + (val - z.x) * zk1;
        z.x = z.x * (1.0 - zk2) + (-val - z.x) * zk2;
        tgladColor += (zk1 + zk2) * s->mb_color_X;
        double zk3 = SGT(z.y, lim), zk4 = SLT(z.y, -lim);
        z.y = z.y * (1.0 - zk3) + (val - z.y) * zk3;
        z.y = z.y * (1.0 - zk4) + (-val - z.y) * zk4;
        tgladColor += (zk3 + zk4) * s->mb_color_Y;
        double zk5 = SGT(z.z, lim), zk6 = SLT(z.z, -lim);
        z.z = z.z * (1.0 - zk5) + (val - z.z) * zk5;
        z.z = z.z * (1.0 - zk6) + (-val - z.z) * zk6;
        tgladColor += (zk5 + zk6) * s->mb_color_Z;
        r = v3len(z);
        double r2 = r * r;
        double tglad_factor2 = fR2 / r2;
        double rk1 = SLT(r2, mR2), rk2 = SLT(r2, fR2);
        double rk21 = (1.0 - rk1) * rk2;
#undef SGT
#undef SLT
        z = v3add(v3mul(z, 1.0 - rk1), v3mul(z, tglad_factor1 * rk1));
        z = v3add(v3mul(z, 1.0 - rk21), v3mul(z, tglad_factor2 * rk21));
        tgladDE = tgladDE * (1.0 - rk1) + tgladDE * (tglad_factor1 * rk1);
        tgladDE = tgladDE * (1.0 - rk21) + tgladDE * (tglad_factor2 * rk21);
        tgladColor += rk1 * s->mb_color_Sp1;
        tgladColor += rk21 * s->mb_color_Sp2;
        z = rot_apply(&f->mb_main, z);
        z = v3add(v3mul(z, scale), constant);
        tgladDE = tgladDE * fabs(scale) + 1.0;
        r = v3len(z);
        break;
      }
      case MBR_FORMULA_AEXION: {  /* :916-947 */
        double ca = s->c_add;
        if (L == 0) {
          double cx = fabs(constant.x + constant.y + constant.z) + ca;
          double cy = fabs(-constant.x - constant.y + constant.z) + ca;
          double cz = fabs(-constant.x + constant.y - constant.z) + ca;
          double cw = fabs(constant.x - constant.y - constant.z) + ca;
          constant = v3m(cx, cy, cz);
          constantw = cw;
          double tx = fabs(z.x + z.y + z.z) + ca;
          double ty = fabs(-z.x - z.y + z.z) + ca;
          double tz = fabs(-z.x + z.y - z.z) + ca;
          double tw = fabs(z.x - z.y - z.z) + ca;
          z = v3m(tx, ty, tz);
          w = tw;
        }
        double tx = z.x * z.x - z.y * z.y + 2.0 * w * z.z + constant.x;
        double ty = z.y * z.y - z.x * z.x + 2.0 * w * z.z + constant.y;
        double tz = z.z * z.z - w * w + 2.0 * z.x * z.y + constant.z;
        double tw = w * w - z.z * z.z + 2.0 * z.x * z.y + constantw;
        z = v3m(tx, ty, tz);
        w = tw;
        r = sqrt(z.x * z.x + z.y * z.y + z.z * z.z + w * w);
        break;
      }
      case MBR_FORMULA_BENESI: {  /* :948-970 */
        double r1 = z.y * z.y + z.z * z.z;
        double newx;
        if (constant.x < 0 || z.x < sqrt(r1)) newx = z.x * z.x - r1;
        else newx = -z.x * z.x + r1;
        r1 = -1.0 / sqrt(r1) * 2.0 * fabs(z.x);
        double newy = r1 * (z.y * z.y - z.z * z.z);
        double newz = r1 * 2.0 * z.y * z.z;
        z.x = newx + constant.x; z.y = newy + constant.y; z.z = newz + constant.z;
        r = v3len(z);
        break;
      }
      case MBR_FORMULA_KALEIDOSCOPIC: {  /* :837-886 */
        if (s->ifs_abs[0]) z.x = fabs(z.x);
        if (s->ifs_abs[1]) z.y = fabs(z.y);
        if (s->ifs_abs[2]) z.z = fabs(z.z);
        for (int i = 0; i < 9; i++) {
          if (!s->ifs_enabled[i]) continue;
          z = rot_apply(&f->ifs_rot[i], z);
          double length = v3dot(z, f->ifs_dir[i]);
          if (length < s->ifs_distance[i])
            z = v3sub(z, v3mul(f->ifs_dir[i], 2.0 * (length - s->ifs_distance[i]) * s->ifs_intensity[i]));
        }
        v3 off = v3from(s->ifs_offset);
        z = v3add(rot_apply(&f->ifs_main, v3sub(z, off)), off);
        if (s->ifs_edge[0] > 0) z.x = s->ifs_edge[0] - fabs(s->ifs_edge[0] - z.x);
        if (s->ifs_edge[1] > 0) z.y = s->ifs_edge[1] - fabs(s->ifs_edge[1] - z.y);
        if (s->ifs_edge[2] > 0) z.z = s->ifs_edge[2] - fabs(s->ifs_edge[2] - z.z);
        if (mode == MODE_COLOURING) { double length2 = v3len(z); if (length2 < min) min = length2; }
        double sc = s->ifs_scale;
        z = v3mul(z, sc);
        if (s->ifs_menger_sponge_mode) {
          z.x -= off.x * (sc - 1.0);
          z.y -= off.y * (sc - 1.0);
          if (z.z > 0.5 * off.z * (sc - 1.0)) z.z -= off.z * (sc - 1.0);
        } else {
          z = v3sub(z, v3mul(off, sc - 1.0));
        }
        tgladDE *= sc;
        r = v3len(z);
        break;
      }
      case MBR_FORMULA_INVERT_X: z.x = z.x >= 0.0 ? z.x * z.x / (fabs(z.x) + p) : -z.x * z.x / (fabs(z.x) + p); r = v3len(z); break;
      case MBR_FORMULA_INVERT_Y: z.y = z.y >= 0.0 ? z.y * z.y / (fabs(z.y) + p) : -z.y * z.y / (fabs(z.y) + p); r = v3len(z); break;
      case MBR_FORMULA_INVERT_Z: z.z = z.z >= 0.0 ? z.z * z.z / (fabs(z.z) + p) : -z.z * z.z / (fabs(z.z) + p); r = v3len(z); break;
      case MBR_FORMULA_INVERT_R: {  /* :1002-1010 */
        double rInv = r * r / (r + p);
        z.x = z.x / r * rInv; z.y = z.y / r * rInv; z.z = z.z / r * rInv;
        r = v3len(z);
        break;
      }
      case MBR_FORMULA_SPHERICAL_FOLD: {  /* :1011-1029 */
        double rr = r * r, pp = p * p;
        if (rr < pp) { z.x = 1.0 / pp; z.y = 1.0 / pp; z.z = 1.0 / pp; }
        else if (rr < pp * 4.0) { z.x = 1.0 / rr; z.y = 1.0 / rr; z.z = 1.0 / rr; }
        r = v3len(z);
        break;
      }
      case MBR_FORMULA_POW_XYZ:  /* :1030-1037 */
        z.x = z.x >= 0 ? pow(z.x, p) : -pow(-z.x, p);
        z.y = z.y >= 0 ? pow(z.y, p) : -pow(-z.y, p);
        z.z = z.z >= 0 ? pow(z.z, p) : -pow(-z.z, p);
        r = v3len(z);
        break;
      case MBR_FORMULA_SCALE_X: z.x = z.x * p; r = v3len(z); break;
      case MBR_FORMULA_SCALE_Y: z.y = z.y * p; r = v3len(z); break;
      case MBR_FORMULA_SCALE_Z: z.z = z.z * p; r = v3len(z); break;
      case MBR_FORMULA_OFFSET_X: z.x = z.x + p; r = v3len(z); break;
      case MBR_FORMULA_OFFSET_Y: z.y = z.y + p; r = v3len(z); break;
      case MBR_FORMULA_OFFSET_Z: z.z = z.z + p; r = v3len(z); break;
      case MBR_FORMULA_ANGLE_MULTIPLY_X: {  /* :1074-1082 */
        double angle = atan2(z.z, z.y) * p, tempR = sqrt(z.z * z.z + z.y * z.y);
        z.y = tempR * cos(angle); z.z = tempR * sin(angle);
        r = v3len(z);
        break;
      }
      case MBR_FORMULA_ANGLE_MULTIPLY_Y: {
        double angle = atan2(z.z, z.x) * p, tempR = sqrt(z.z * z.z + z.x * z.x);
        z.x = tempR * cos(angle); z.z = tempR * sin(angle);
        r = v3len(z);
        break;
      }
      case MBR_FORMULA_ANGLE_MULTIPLY_Z: {
        double angle = atan2(z.y, z.x) * p, tempR = sqrt(z.x * z.x + z.y * z.y);
        z.x = tempR * cos(angle); z.y = tempR * sin(angle);
        r = v3len(z);
        break;
      }
      case MBR_FORMULA_BRISTORBROT: {  /* :971-983 */
        double newx = z.x * z.x - z.y * z.y - z.z * z.z;
        double newy = z.y * (2.0 * z.x - z.z);
        double newz = z.z * (2.0 * z.x + z.y);
        z.x = newx + constant.x; z.y = newy + constant.y; z.z = newz + constant.z;
        r = v3len(z);
        break;
      }
      case MBR_FORMULA_XENODREAMBUIE: {
        double rp = pow(r, p);
        double th = atan2(z.y, z.x);
        double ph = acos(z.z / r);
        if (ph > 0.5 * M_PI) ph = M_PI - ph;
        else if (ph < -0.5 * M_PI) ph = -M_PI - ph;
        z.x = rp * cos(th * p) * sin(ph * p);
        z.y = rp * sin(th * p) * sin(ph * p);
        z.z = rp * cos(ph * p);
        z = v3add(z, constant);
        r = v3len(z);
        break;
      }
      default:
        break;
    }

    if (mode == MODE_DELTA1) {
      if (r > 1e10) break;
    } else if (mode == MODE_DELTA2) {
      if (L == delta_limit) break;
    }

    /* the fake lights' orbit trap (:1121-1130): inverse squared distances to
     * the trap point over iterations min_iter..max_iter; the sum is returned
     * when the orbit leaves the trap's 1000 radius, else the last distance
     * (or, as in every mode, the escape value of the Menger and Mandelbox
     * groups below) */
    if (mode == MODE_ORBIT_TRAP) {
      distance = v3len(v3sub(z, v3from(s->fake_lights_orbit_trap)));
      if (L >= s->fake_lights_min_iter && L <= s->fake_lights_max_iter) orbitTrapTotal += (1.0 / (distance * distance));
      if (distance > 1000) { distance = orbitTrapTotal; break; }
    }

    if (formula == MBR_FORMULA_MENGER_SPONGE || formula == MBR_FORMULA_KALEIDOSCOPIC) {
      if (r > 1000) { distance = (r - 2.0) / tgladDE; break; }
    } else if (formula == MBR_FORMULA_TGLAD || formula == MBR_FORMULA_SMOOTH_MANDELBOX ||
               formula == MBR_FORMULA_MANDELBOX_VARY_SCALE_4D || formula == MBR_FORMULA_GENERALIZED_FOLD_BOX) {
      if (r > 1024) { distance = r / fabs(tgladDE); break; }
    } else {
      if (mode == MODE_NORMAL) {
        if (r > s->bailout) { distance = 0.5 * r * log(r) / r_dz; break; }
      } else if (mode == MODE_COLOURING) {
        if (s->primitive_only_plane) {
          distance = v3len(z);
          if (distance > 1e15) { distance = (L - log(log(r) / log((double)N)) / log(p)) / 100.0; break; }
        } else {
          distance = v3len(z);
          if (distance < min) min = distance;
          if (distance > 1e15) { distance = min; break; }
        }
      }
    }
  }

  if (iter_count) *iter_count = L;
  if (mode == MODE_NORMAL) {
    if (L == N) distance = 0;
    return distance;
  }
  if (mode == MODE_DELTA1 || mode == MODE_DELTA2) return r;
  if (mode == MODE_ORBIT_TRAP) return distance;
  if (s->formula == MBR_FORMULA_HYBRID) {  /* src/fractal.cpp:1218-1229 */
    if (min > 100) min = 100;
    if (distance > 20) distance = 20;
    if (tgladColor > 1000) tgladColor = 1000;
    return distance * 5000.0 + tgladColor * 100.0 + min * 1000.0;
  }
  if (formula == MBR_FORMULA_TGLAD || formula == MBR_FORMULA_SMOOTH_MANDELBOX ||
      formula == MBR_FORMULA_MANDELBOX_VARY_SCALE_4D || formula == MBR_FORMULA_GENERALIZED_FOLD_BOX)
    return tgladColor * 100.0 + v3len(z) * s->mb_color_R;
  if (formula == MBR_FORMULA_MENGER_SPONGE || formula == MBR_FORMULA_KALEIDOSCOPIC) return min * 1000.0;
  return distance * 5000.0;
}

/* -------------------------------------------------------------- primitives */
/* src/primitives.cpp:15-119 */
static double prim_plane(v3 p, v3 centre, v3 normal) { return v3dot(v3norm(normal), v3sub(p, centre)); }

static double prim_box(v3 p, v3 c, v3 size, int inverted) {
  v3 c1 = v3m(c.x - 0.5 * size.x, c.y - 0.5 * size.y, c.z - 0.5 * size.z);
  v3 c2 = v3m(c.x + 0.5 * size.x, c.y + 0.5 * size.y, c.z + 0.5 * size.z);
  double sg = inverted ? 1.0 : -1.0;
  double d[6];
  d[0] = prim_plane(p, c1, v3m(sg, 0, 0));
  d[1] = prim_plane(p, c2, v3m(-sg, 0, 0));
  d[2] = prim_plane(p, c1, v3m(0, sg, 0));
  d[3] = prim_plane(p, c2, v3m(0, -sg, 0));
  d[4] = prim_plane(p, c1, v3m(0, 0, sg));
  d[5] = prim_plane(p, c2, v3m(0, 0, -sg));
  double dist = d[0];
  for (int i = 1; i < 6; i++) {
    if (inverted) dist = (d[i] < dist) ? d[i] : dist;
    else dist = (d[i] > dist) ? d[i] : dist;
  }
  return dist;
}

static double prim_water(v3 p, double height, double amplitude, double length, double rotation,
                         int iterations, double animSpeed, int frame) {
  v3 plane = v3norm(v3m(0, 0, -1));
  double planeDistance = v3dot(plane, v3sub(p, v3m(0, 0, height)));
  if (planeDistance < amplitude * 10.0) {
    rotm rm; rot_null(&rm); rot_z(&rm, rotation / 180 * M_PI);
    p = rot_apply(&rm, p);
    double phase = animSpeed * frame;
    double k = 0.23, waveX = 0, waveY = 0, pp = 1.0, p2 = 0.05;
    for (int i = 1; i <= iterations; i++) {
      float p3 = (float)(pp * p2);
      double shift = phase / (i / 3.0 + 1.0);
      double wx = sin(i + 0.4 * (waveX) * p3 + sin(k * p.y / length * p3) + p.x / length * p3 + shift) / pp;
      double wy = cos(i + 0.4 * (waveY) * p3 + sin(p.x / length * p3) + k * p.y / length * p3 + shift * 0.23) / pp;
      waveX += wx; waveY += wy;
      p2 = p2 + (1.0 - p2) * 0.7;
      pp *= 1.872;
    }
    planeDistance += (waveX + waveY) * amplitude;
  }
  return planeDistance;
}

/* ------------------------------------------------------ CalculateDistance */
typedef struct { uint64_t calls, iterations; } de_counter;

/* CalculateDistance (src/fractal.cpp:1245-1448), analytic-DE branch (the
 * supported formulas 2, 7 and 8 are analytic, src/interface.cpp:690-693). */
static double calc_distance(const fractal_ctx *f, v3 point, double detailSize, int N, int minN,
                            int normalMode, int *max_iter, int *iters_out, int *object_out,
                            de_counter *cnt) {
  const mbr_scene *s = f->s;
  int L = 0, object = MBR_OBJ_FRACTAL, mi = 0;
  double distance;
  double limitBoxDist = 0.0;

  if (s->limits_enabled) {
    double da = fmax(point.x - s->limit_max[0], -(point.x - s->limit_min[0]));
    double db = fmax(point.y - s->limit_max[1], -(point.y - s->limit_min[1]));
    double dc = fmax(point.z - s->limit_max[2], -(point.z - s->limit_min[2]));
    /* dMax (src/common_math.cpp:34-44) */
    limitBoxDist = (da > db) ? ((da > dc) ? da : dc) : ((db > dc) ? db : dc);
    if (limitBoxDist > detailSize) {
      if (max_iter) *max_iter = 0;
      if (object_out) *object_out = object;
      return limitBoxDist;
    }
  }

  const int analytic = s->formula == MBR_FORMULA_TRIG_DE || s->formula == MBR_FORMULA_TRIG_OPTIM ||
                       s->formula == MBR_FORMULA_MENGER_SPONGE || s->formula == MBR_FORMULA_TGLAD ||
                       s->formula == MBR_FORMULA_SMOOTH_MANDELBOX || s->formula == MBR_FORMULA_KALEIDOSCOPIC ||
                       s->formula == MBR_FORMULA_MANDELBOX_VARY_SCALE_4D ||
                       s->formula == MBR_FORMULA_GENERALIZED_FOLD_BOX;
  if (!s->primitive_only_plane && !analytic) {
    /* delta DE (src/fractal.cpp:1307-1342): orbit radius at the point and
     * at +1e-10 along x, y, z, the point updated in place as there */
    const double deltaDE = 1e-10;
    double r = compute(f, MODE_DELTA1, point, N, &L);
    const int retval = L;
    if (iters_out) *iters_out = L;
    point.x += deltaDE;
    double r2 = compute(f, MODE_DELTA2, point, N, &L);
    double dr1 = fabs(r2 - r) / deltaDE;
    point.x -= deltaDE;
    point.y += deltaDE;
    r2 = compute(f, MODE_DELTA2, point, N, &L);
    double dr2 = fabs(r2 - r) / deltaDE;
    point.y -= deltaDE;
    point.z += deltaDE;
    r2 = compute(f, MODE_DELTA2, point, N, &L);
    double dr3 = fabs(r2 - r) / deltaDE;
    point.z -= deltaDE;  /* (the reference's point is a copy: the caller's point is unchanged) */
    double dr = sqrt(dr1 * dr1 + dr2 * dr2 + dr3 * dr3);
    distance = s->linear_DE_mode ? 0.5 * r / dr : 0.5 * r * log(r) / dr;
    if (cnt) { cnt->calls++; cnt->iterations += 4 * ((retval < N) ? (uint64_t)retval + 1 : (uint64_t)N); }
    if (distance < 0) distance = 0;
    mi = 0;
    if (retval == N) { mi = 1; distance = 0; }
    if (L < minN && distance < detailSize) distance = detailSize;
    if (s->interior_mode && !normalMode) {
      if (distance < 0.5 * detailSize || L == N) { distance = detailSize; mi = 0; }
    } else if (s->interior_mode && normalMode) {
      if (distance < 0.9 * detailSize) { distance = detailSize - distance; mi = 0; }
    }
    if (s->iter_thresh && retval < N) {
      if (distance < detailSize) distance = detailSize * 1.01;
    }
  } else if (!s->primitive_only_plane) {
    distance = compute(f, MODE_NORMAL, point, N, &L);
    if (cnt) { cnt->calls++; cnt->iterations += (L < N) ? (uint64_t)L + 1 : (uint64_t)N; }
    mi = (L == N);
    if (iters_out) *iters_out = L;
    if (distance < 0) distance = 0;
    if (L < minN && distance < detailSize) distance = detailSize;
    if (s->interior_mode && !normalMode) {
      if (distance < 0.5 * detailSize || L == N) { distance = detailSize; mi = 0; }
    } else if (s->interior_mode && normalMode) {
      if (distance < 0.9 * detailSize) { distance = detailSize - distance; mi = 0; }
    }
    if (s->iter_thresh && L < N) {
      if (distance < detailSize) distance = detailSize * 1.01;
    }
  } else {
    distance = 10.0;
    mi = 0;
  }

  if (s->plane_enabled) {
    double d = prim_plane(point, v3from(s->plane_centre), v3from(s->plane_normal));
    if (!s->primitive_only_plane && d < distance) object = MBR_OBJ_PLANE;
    distance = (d < distance) ? d : distance;
  }
  if (s->box_enabled) {
    double d = prim_box(point, v3from(s->box_centre), v3from(s->box_size), 0);
    if (d < distance) object = MBR_OBJ_BOX;
    distance = (d < distance) ? d : distance;
  }
  if (s->inv_box_enabled) {
    double d = prim_box(point, v3from(s->inv_box_centre), v3from(s->inv_box_size), 1);
    if (d < distance) object = MBR_OBJ_BOX_INV;
    distance = (d < distance) ? d : distance;
  }
  if (s->sphere_enabled) {
    double d = v3len(v3sub(point, v3from(s->sphere_centre))) - s->sphere_radius;
    if (d < distance) object = MBR_OBJ_SPHERE;
    distance = (d < distance) ? d : distance;
  }
  if (s->inv_sphere_enabled) {
    double d = s->inv_sphere_radius - v3len(v3sub(point, v3from(s->inv_sphere_centre)));
    if (d < distance) object = MBR_OBJ_SPHERE_INV;
    distance = (d < distance) ? d : distance;
  }
  if (s->water_enabled) {
    double d = prim_water(point, s->water_height, s->water_amplitude, s->water_length,
                          s->water_rotation, s->water_iterations, s->water_anim_speed, s->frame_no);
    if (d < distance) object = MBR_OBJ_WATER;
    distance = (d < distance) ? d : distance;
  }
  if (s->limits_enabled) {
    if (limitBoxDist < detailSize) distance = fmax(distance, limitBoxDist);
  }
  if (max_iter) *max_iter = mi;
  if (object_out) *object_out = object;
  return distance;
}

/* ---------------------------------------------------------------- camera */
typedef struct {
  rotm R;
  v3 start, light;
  double resolution;
  int aspect_w, aspect_h;
} camera_t;

/* MainThread set-up (src/Render3D.cpp:236-256, :316-333) plus the
 * resolution of ReadInterface (src/interface.cpp:695). */
static void camera_init(camera_t *c, const mbr_scene *s) {
  rot_null(&c->R);
  rot_z(&c->R, s->alpha);
  rot_x(&c->R, s->beta);
  rot_y(&c->R, s->gamma);
  v3 baseY = rot_apply(&c->R, v3m(0.0, 1.0, 0.0));
  v3 vec;
  vec.x = cos(s->main_light_alpha - 0.5 * M_PI) * cos(-s->main_light_beta);
  vec.y = sin(s->main_light_alpha - 0.5 * M_PI) * cos(-s->main_light_beta);
  vec.z = sin(-s->main_light_beta);
  c->light = rot_apply(&c->R, vec);
  if (s->perspective_type == MBR_PERSP_FISH_EYE || s->perspective_type == MBR_PERSP_EQUIRECTANGULAR)
    c->start = v3from(s->vp);
  else
    c->start = v3sub(v3from(s->vp), v3mul(baseY, 1.0 / s->persp * s->zoom));
  c->resolution = 1.0 / s->image_width / s->tiles;
}

/* per-pixel view vector (src/Render3D.cpp:389-438) */
static int view_vector(const mbr_scene *s, const camera_t *c, int x, int y, v3 *out) {
  int width = s->image_width, height = s->image_height, tiles = s->tiles;
  double aspectRatio = (double)width / height;
  double tileX = (s->tile_index % tiles), tileY = (s->tile_index / tiles);
  double fov = s->persp;
  double x2 = ((double)x / width / tiles - 0.5 + tileX / tiles) * aspectRatio;
  double z2 = ((double)y / height / tiles - 0.5 + tileY / tiles);
  int cut = 0;
  v3 v;
  if (s->perspective_type == MBR_PERSP_FISH_EYE) {
    if (s->fish_eye_cut && sqrt(x2 * x2 + z2 * z2) > 0.5 / fov) cut = 1;
    double x3 = x2 * M_PI, z3 = z2 * M_PI;
    double r = sqrt(x3 * x3 + z3 * z3);
    if (r == 0) v = v3m(0.0, 1.0, 0.0);
    else v = v3m(x3 / r * sin(r * fov), cos(r * fov), z3 / r * sin(r * fov));
    v = v3norm(v);
  } else if (s->perspective_type == MBR_PERSP_EQUIRECTANGULAR) {
    double x3 = x2 * M_PI, z3 = z2 * M_PI;
    v = v3m(sin(fov * x3) * cos(fov * z3), cos(fov * x3) * cos(fov * z3), sin(fov * z3));
    v = v3norm(v);
  } else {
    v = v3m(x2 * fov, 1.0, z2 * fov);
  }
  *out = rot_apply(&c->R, v);
  return cut;
}

/* ---------------------------------------------------------------- march */
typedef struct { double distance, step, distThresh; v3 point; int iters; } step_rec;

typedef struct {
  int mode;               /* enum oracle_jitter */
  uint32_t seed, frame, pixel;
} jitter_t;

/* Random(1000) (src/common_math.cpp:25-32) or its hash replacement */
static inline int jitter_draw(const jitter_t *j, int step) {
  if (j->mode == ORACLE_JITTER_GLIBC) return rand() % 1001;
  if (j->mode == ORACLE_JITTER_HASH) return (int)oracle_jitter_draw(j->seed, j->frame, j->pixel, (uint32_t)step);
  return 0;
}

typedef struct {
  v3 point;
  double depth, distThresh, lastDist, detail;
  int found, buffCount, object, de_calls, missed;
} march_result;

/* RayMarching (src/Render3D.cpp:79-199) with distThreshInit = 0 (primary
 * rays, :444) and binary search enabled. */
static void ray_march(const fractal_ctx *f, const camera_t *c, v3 start, v3 direction, double maxScan,
                      double thr_init, const jitter_t *jit, step_rec *stepBuff, march_result *out,
                      de_counter *cnt) {
  const mbr_scene *s = f->s;
  v3 point = v3m(0, 0, 0);
  int found = 0;
  double scan = s->view_distance_min;
  double distThresh = 0.0;
  double dist = 0;
  double search_accuracy = 0.01 * s->quality;
  double search_limit = 1.0 - search_accuracy;
  int counter = 0;
  double step = s->view_distance_min;
  int buffCount = 0;
  double distThreshInit = thr_init;  /* *distThreshOut on entry: 0, or the previous ray's (reflections) */
  double detail = 0.0;
  int object = MBR_OBJ_FRACTAL;
  int fisheye = (s->perspective_type == MBR_PERSP_FISH_EYE || s->perspective_type == MBR_PERSP_EQUIRECTANGULAR);
  out->missed = 0;

  for (int i = 0; i < 10000; i++) {
    counter++;
    point = v3add(start, v3mul(direction, scan));
    int max_iter = 0, iters = 0;
    if (s->constant_DE_threshold) distThresh = s->quality;
    else distThresh = scan * c->resolution * s->persp / s->quality + distThreshInit;
    if (fisheye) distThresh *= M_PI;
    detail = distThresh;
    dist = calc_distance(f, point, detail, s->N, s->minN, 0, &max_iter, &iters, &object, cnt);
    stepBuff[i].distance = dist;
    stepBuff[i].iters = iters;
    stepBuff[i].distThresh = distThresh;
    if (dist > 3.0) dist = 3.0;
    if (dist < distThresh) {
      if (dist < 0.1 * distThresh) out->missed = 1;
      found = 1;
      break;
    }
    stepBuff[i].step = step;
    double jf = 1.0 - jitter_draw(jit, i) / 10000.0;
    if (s->interior_mode) step = (dist - 0.8 * distThresh) * s->DE_factor * jf;
    else step = (dist - 0.5 * distThresh) * s->DE_factor * jf;
    stepBuff[i].point = point;
    buffCount++;
    scan += step;
    if (scan > maxScan) break;
  }

  if (found) {
    step *= 0.5;
    for (int i = 0; i < 20; i++) {
      if (dist < distThresh && dist > distThresh * search_limit) break;
      if (dist > distThresh) { scan += step; point = v3add(start, v3mul(direction, scan)); }
      else if (dist < distThresh * search_limit) { scan -= step; point = v3add(start, v3mul(direction, scan)); }
      if (s->constant_DE_threshold) distThresh = s->quality;
      else distThresh = scan * c->resolution * s->persp / s->quality + distThreshInit;
      int max_iter = 0;
      dist = calc_distance(f, point, detail, s->N, s->minN, 0, &max_iter, NULL, &object, cnt);
      counter++;
      step *= 0.5;
    }
  }
  out->point = point;
  out->depth = scan;
  out->distThresh = distThresh;
  out->lastDist = dist;
  out->found = found;
  out->buffCount = buffCount;
  out->object = object;
  out->de_calls = counter;
  out->detail = detail;
}

/* ---------------------------------------------------------------- shaders */
typedef struct { double R, G, B, alpha; } shade_t;

typedef struct {
  const fractal_ctx *f;
  const camera_t *c;
  v3 point, viewVector, normal, lightVect;
  double dist_thresh, lastDist, delta, depth;
  double detail;          /* calcParam->doubles.detailSize as the reference mutates it */
  int objectType;
  const double *ao_dirs; const int *ao_rgb; int ao_count;
  const struct light_set *lights;
  de_counter *cnt;
} shader_in;

/* The Lights array and lightsPlaced of src/shaders.cpp:17-18: light i takes
 * part when i < auxLightNumber or it is enabled (:1059, :248). */
typedef struct light_set {
  const mbr_light *tab;
  int number;   /* auxLightNumber */
  int count;    /* lightsPlaced, at least 4 */
  int any;      /* some light takes part */
} light_set;

static int light_used(const light_set *L, int i) { return i < L->number || L->tab[i].enabled; }

/* CalculateNormals (src/shaders.cpp:675-757) */
static v3 calculate_normals(shader_in *in) {
  const mbr_scene *s = in->f->s;
  v3 normal = v3m(0, 0, 0);
  int mi;
  if (!s->slow_shading) {
    double delta = in->delta * s->smoothness;
    if (s->interior_mode) delta = in->dist_thresh * s->quality * 0.2 * s->smoothness;
    int N5 = s->N * 5;
    double s1 = calc_distance(in->f, in->point, in->detail, N5, 0, 1, &mi, NULL, NULL, in->cnt);
    double s2 = calc_distance(in->f, v3add(in->point, v3m(delta, 0.0, 0.0)), in->detail, N5, 0, 1, &mi, NULL, NULL, in->cnt);
    double s3 = calc_distance(in->f, v3add(in->point, v3m(0.0, delta, 0.0)), in->detail, N5, 0, 1, &mi, NULL, NULL, in->cnt);
    double s4 = calc_distance(in->f, v3add(in->point, v3m(0.0, 0.0, delta)), in->detail, N5, 0, 1, &mi, NULL, NULL, in->cnt);
    normal.x = s2 - s1; normal.y = s3 - s1; normal.z = s4 - s1;
  } else {
    double delta = in->delta * s->smoothness;
    int N5 = s->N * 5;
    if (s->interior_mode) delta = in->dist_thresh * s->quality * 0.2 * s->smoothness;
    v3 p2;
    for (p2.x = -1.0; p2.x <= 1.0; p2.x += 0.2)
      for (p2.y = -1.0; p2.y <= 1.0; p2.y += 0.2)
        for (p2.z = -1.0; p2.z <= 1.0; p2.z += 0.2) {
          v3 p3 = v3add(in->point, v3mul(p2, delta));
          double dist = calc_distance(in->f, p3, in->detail, N5, s->minN, 1, &mi, NULL, NULL, in->cnt);
          normal = v3add(normal, v3mul(p2, dist));
        }
  }
  if (normal.x == 0 && normal.y == 0 && normal.z == 0) normal.x = 1;
  else normal = v3norm(normal);
  return normal;
}

/* imageSwitches.volumetricLightEnabled (src/interface.cpp:697-702) or
 * iteration fog: the shadow marches then take unit DE steps and a threshold
 * from the distance to the eye (src/shaders.cpp:494, :511-519, :1084, :1102-1110) */
static int shadow_fx(const mbr_scene *s) {
  int v = s->iter_fog_enabled;
  for (int i = 0; i < 5; i++)
    if (s->volumetric_light_enabled[i]) v = 1;
  return v;
}

static float fx_thresh(const shader_in *in, v3 point2) {
  const mbr_scene *s = in->f->s;
  if (!s->constant_DE_threshold)
    return (float)(v3len(v3sub(point2, in->c->start)) * in->c->resolution * s->persp / s->quality);
  return (float)s->quality;
}

/* IterOpacity (src/shaders.cpp:1214-1222) */
static double iter_opacity(double step, double iters, double maxN, double trim, double opacitySp) {
  double opacity = (iters - trim) / maxN;
  if (opacity < 0.0) opacity = 0.0;
  opacity *= opacity;
  opacity *= step * opacitySp;
  if (opacity > 1.0) opacity = 1.0;
  return opacity;
}

/* MainShadow (src/shaders.cpp:480-570) */
static double main_shadow(shader_in *in) {
  const mbr_scene *s = in->f->s;
  const int fx = shadow_fx(s);
  double factor = in->delta / in->c->resolution;
  if (!s->penetrating_lights) factor = s->view_distance_max;
  double dist = in->dist_thresh;
  in->detail = in->dist_thresh;
  double DE_factor = fx ? 1.0 : s->DE_factor;
  double start = in->dist_thresh;
  if (s->interior_mode) start = in->dist_thresh * DE_factor;
  double shadowTemp = 1.0;
  double softRange = tan(s->shadow_cone_angle / 180.0 * M_PI);
  double maxSoft = 0.0;
  int bSoft = (!s->iter_fog_enabled && !s->limits_enabled && !s->iter_thresh) && softRange > 0.0;
  int mi, it;
  for (double i = start; i < factor; i += dist * DE_factor) {
    v3 point2 = v3add(in->point, v3mul(in->lightVect, i));
    float dist_thresh = fx ? fx_thresh(in, point2) : (float)in->dist_thresh;
    in->detail = dist_thresh;
    dist = calc_distance(in->f, point2, in->detail, s->N, s->minN, 0, &mi, &it, NULL, in->cnt);
    if (bSoft) {
      double angle = (dist - dist_thresh) / i;
      if (angle < 0) angle = 0;
      if (dist < dist_thresh) angle = 0;
      double softShadow = (1.0 - angle / softRange);
      if (s->penetrating_lights) softShadow *= (factor - i) / factor;
      if (softShadow < 0) softShadow = 0;
      if (softShadow > maxSoft) maxSoft = softShadow;
    }
    double opacity = s->iter_fog_enabled
                         ? iter_opacity(dist * DE_factor, it, s->N, s->iter_fog_opacity_trim, s->iter_fog_opacity)
                         : 0.0;
    shadowTemp -= opacity * (factor - i) / factor;
    if (dist < dist_thresh || shadowTemp < 0.0) {
      shadowTemp -= (factor - i) / factor;
      if (!s->penetrating_lights) shadowTemp = 0.0;
      if (shadowTemp < 0.0) shadowTemp = 0.0;
      break;
    }
  }
  return bSoft ? (1.0 - maxSoft) : shadowTemp;
}

/* FastAmbientOcclusion2 (src/shaders.cpp:585-605) */
static double fast_ao2(shader_in *in) {
  const mbr_scene *s = in->f->s;
  double delta = in->dist_thresh;
  double aoTemp = 0;
  double quality = s->ao_quality;
  int mi;
  for (int i = 1; i < quality * quality; i++) {
    double scan = i * i * delta;
    v3 pt = v3add(in->point, v3mul(in->normal, scan));
    double dist = calc_distance(in->f, pt, in->detail, s->N, s->minN, 0, &mi, NULL, NULL, in->cnt);
    aoTemp += 1.0 / (pow(2.0, i)) * (scan - s->fast_ao_tune * dist) / in->dist_thresh;
  }
  double ao = 1.0 - 0.2 * aoTemp;
  if (ao < 0) ao = 0;
  return ao;
}

/* AmbientOcclusion (src/shaders.cpp:607-673) */
static shade_t ambient_occlusion(shader_in *in) {
  const mbr_scene *s = in->f->s;
  shade_t AO = {0, 0, 0, 0};
  int max_iter;
  double start_dist = in->delta;
  double end_dist = in->delta / in->c->resolution;
  for (int k = 0; k < in->ao_count; k++) {
    v3 v = v3from(in->ao_dirs + 3 * k);
    double dist = in->lastDist;
    double shadowTemp = 1.0;
    for (double r = start_dist; r < end_dist; r += dist * 2.0) {
      v3 point2 = v3add(in->point, v3mul(v, r));
      dist = calc_distance(in->f, point2, in->detail, s->N, s->minN, 0, &max_iter, NULL, NULL, in->cnt);
      float dist_thresh = (float)in->dist_thresh;
      if (dist < dist_thresh || max_iter || shadowTemp < 0.0) {
        shadowTemp -= (end_dist - r) / end_dist;
        if (shadowTemp < 0.0) shadowTemp = 0.0;
        break;
      }
    }
    AO.R += shadowTemp * in->ao_rgb[3 * k + 0];
    AO.G += shadowTemp * in->ao_rgb[3 * k + 1];
    AO.B += shadowTemp * in->ao_rgb[3 * k + 2];
  }
  AO.R /= (in->ao_count * 256.0);
  AO.G /= (in->ao_count * 256.0);
  AO.B /= (in->ao_count * 256.0);
  return AO;
}

/* IndexToColour (src/shaders.cpp:1273-1305) */
static mbr_rgb index_to_colour(int index, const mbr_rgb *palette) {
  mbr_rgb colour = {255, 255, 255};
  if (index < 0) return palette[255];
  index = index % (255 * 256);
  int kol = index / 256;
  if (kol < 255) {
    double R1 = palette[kol].R, G1 = palette[kol].G, B1 = palette[kol].B;
    double R2 = palette[kol + 1].R, G2 = palette[kol + 1].G, B2 = palette[kol + 1].B;
    double RK = (R2 - R1) / 256.0, GK = (G2 - G1) / 256.0, BK = (B2 - B1) / 256.0;
    int delta = index % 256;
    colour.R = (int)(R1 + (RK * delta));
    colour.G = (int)(G1 + (GK * delta));
    colour.B = (int)(B1 + (BK * delta));
  }
  return colour;
}

/* SurfaceColour (src/shaders.cpp:813-926) */
static shade_t surface_colour(shader_in *in) {
  const mbr_scene *s = in->f->s;
  shade_t out = {0, 0, 0, 0};
  mbr_rgb pc;
  switch (in->objectType) {
    case MBR_OBJ_FRACTAL: {
      int it;
      double cval = compute(in->f, MODE_COLOURING, in->point, s->N * 10, &it);
      if (in->cnt) { in->cnt->calls++; in->cnt->iterations += (it < s->N * 10) ? (uint64_t)it + 1 : (uint64_t)s->N * 10; }
      int nrCol = x86_d2i(floor(cval));
      nrCol = x86_abs(nrCol) % (248 * 256);
      mbr_rgb colour = {256, 256, 256};
      if (s->coloring_enabled) {
        int color_number;
        if (nrCol >= 248 * 256) color_number = nrCol;
        else color_number = x86_d2i(nrCol * s->coloring_speed + 256 * s->palette_offset) % 65536;
        colour = index_to_colour(color_number, s->palette);
      }
      out.R = colour.R / 256.0; out.G = colour.G / 256.0; out.B = colour.B / 256.0;
      return out;
    }
    case MBR_OBJ_WATER: pc = s->water_colour; break;
    case MBR_OBJ_BOX: pc = s->box_colour; break;
    case MBR_OBJ_BOX_INV: pc = s->inv_box_colour; break;
    case MBR_OBJ_SPHERE: pc = s->sphere_colour; break;
    case MBR_OBJ_SPHERE_INV: pc = s->inv_sphere_colour; break;
    default: pc = s->plane_colour; break;
  }
  out.R = pc.R / 65536.0; out.G = pc.G / 65536.0; out.B = pc.B / 65536.0;
  return out;
}

/* AuxShadow (src/shaders.cpp:1075-1132) */
static double aux_shadow(shader_in *in, double distance, v3 lightVector) {
  const mbr_scene *s = in->f->s;
  const int fx = shadow_fx(s);
  double dist = in->delta;
  double shadowTemp = 1.0;
  double DE_factor = fx ? 1.0 : s->DE_factor;
  int mi, it;
  for (double i = in->delta; i < distance; i += dist * DE_factor) {
    v3 point2 = v3add(in->point, v3mul(lightVector, i));
    dist = calc_distance(in->f, point2, in->detail, s->N, s->minN, 0, &mi, &it, NULL, in->cnt);
    double opacity = s->iter_fog_enabled
                         ? iter_opacity(dist * DE_factor, it, s->N, s->iter_fog_opacity_trim, s->iter_fog_opacity)
                         : 0.0;
    shadowTemp -= opacity * (distance - i) / distance;
    float dist_thresh = fx ? fx_thresh(in, point2) : (float)in->dist_thresh;
    if (dist < dist_thresh || mi || shadowTemp < 0.0) {
      if (s->penetrating_lights) {
        shadowTemp -= (distance - i) / distance;
        if (shadowTemp < 0.0) shadowTemp = 0.0;
      } else {
        shadowTemp = 0.0;
      }
      break;
    }
  }
  return shadowTemp;
}

/* LightShading (src/shaders.cpp:997-1049) */
static shade_t light_shading(shader_in *in, const mbr_light *light, int number, shade_t *outSpecular) {
  const mbr_scene *s = in->f->s;
  v3 d = v3sub(v3from(light->position), in->point);
  double distance = v3len(d);
  v3 lightVector = v3norm(d);
  double intensity = s->aux_light_intensity * 100.0 * light->intensity / (distance * distance) / number;
  double shade = v3dot(in->normal, lightVector);
  if (shade < 0) shade = 0;
  shade = shade * intensity;
  if (shade > 500.0) shade = 500.0;
  v3 half = v3norm(v3sub(lightVector, in->viewVector));
  double shade2 = v3dot(in->normal, half);
  if (shade2 < 0.0) shade2 = 0.0;
  shade2 = pow(shade2, 30.0) * 1.0;
  shade2 *= intensity * s->specular;
  if (shade2 > 15.0) shade2 = 15.0;
  if ((shade > 0.01 || shade2 > 0.01) && s->shadow) {
    double l = aux_shadow(in, distance, lightVector);
    shade *= l;
    shade2 *= l;
  } else if (s->shadow) {
    shade = 0;
    shade2 = 0;
  }
  shade_t out = {shade * light->colour[0] / 65536.0, shade * light->colour[1] / 65536.0,
                 shade * light->colour[2] / 65536.0, 0.0};
  outSpecular->R = shade2 * light->colour[0] / 65536.0;
  outSpecular->G = shade2 * light->colour[1] / 65536.0;
  outSpecular->B = shade2 * light->colour[2] / 65536.0;
  return out;
}

/* AuxLightsShader (src/shaders.cpp:1051-1073) */
static shade_t aux_lights_shader(shader_in *in, shade_t *specularOut) {
  shade_t sum = {0, 0, 0, 0}, spec = {0, 0, 0, 0};
  const light_set *L = in->lights;
  if (L && L->any) {
    for (int i = 0; i < L->count; i++) {
      if (!light_used(L, i)) continue;
      shade_t sp;
      shade_t sh = light_shading(in, &L->tab[i], L->count, &sp);
      sum.R += sh.R; sum.G += sh.G; sum.B += sh.B;
      spec.R += sp.R; spec.G += sp.G; spec.B += sp.B;
    }
  }
  *specularOut = spec;
  return sum;
}

/* Random (src/common_math.cpp:25-32) as built on Linux */
static int ref_random(int max) { return rand() % (max + 1); }

static double dmax3(double a, double b, double c) {
  if (a > b) return a > c ? a : c;
  return b > c ? b : c;
}

/* PlaceRandomLights(params, false) (src/shaders.cpp:1134-1212).  The two
 * draws of one coordinate are taken left to right (the order GCC gives the
 * reference's expression).  Returns aux_light_number + 4 records, -1 when
 * `capacity` is short, -2 after 20 million draws without a placement. */
int oracle_place_lights(const mbr_scene *s, mbr_light *lights, int capacity) {
  int n = s->aux_light_number > 0 ? s->aux_light_number : 0;
  if (capacity < n + 4) return -1;
  memset(lights, 0, sizeof(mbr_light) * (size_t)(n + 4));
  fractal_ctx f; fractal_ctx_init(&f, s);
  srand((unsigned)s->aux_light_random_seed);
  int trial_number = 0;
  double radius_multiplier = 1.0;
  long long draws = 0;
  v3 centre = v3from(s->aux_light_random_center);
  for (int i = 0; i < n; i++) {
    trial_number++;
    if (++draws > 20000000) return -2;
    double c[3];
    for (int k = 0; k < 3; k++) {
      double a = ref_random(2000000) / 1000000.0 - 1.0;
      double b = ref_random(1000000) / 1.0e12;
      c[k] = a + b;
    }
    v3 position = v3add(centre, v3mul(v3mul(v3from(c), s->aux_light_distribution_radius), radius_multiplier));
    int mi = 0;
    double distance = calc_distance(&f, position, 0.0, s->N, s->minN, 0, &mi, NULL, NULL, NULL);
    if (trial_number > 1000) {
      radius_multiplier *= 1.1;
      trial_number = 0;
    }
    if (distance > 0 && !mi && distance < s->aux_light_max_dist * radius_multiplier) {
      radius_multiplier = 1.0;
      lights[i].position[0] = position.x; lights[i].position[1] = position.y; lights[i].position[2] = position.z;
      int col[3];
      col[0] = 20000 + ref_random(80000);
      col[1] = 20000 + ref_random(80000);
      col[2] = 20000 + ref_random(80000);
      double maxColour = dmax3(col[0], col[1], col[2]);
      for (int k = 0; k < 3; k++) lights[i].colour[k] = (int)(col[k] * 65536.0 / maxColour);
      lights[i].intensity = distance * distance / s->aux_light_max_dist;
      lights[i].enabled = 1;
    } else {
      i--;
    }
  }
  for (int i = 0; i < 4; i++) {
    if (s->aux_light_pre_enabled[i]) {
      for (int k = 0; k < 3; k++) lights[i].position[k] = s->aux_light_pre[i][k];
      lights[i].intensity = s->aux_light_pre_intensity[i];
      lights[i].colour[0] = s->aux_light_pre_colour[i].R;
      lights[i].colour[1] = s->aux_light_pre_colour[i].G;
      lights[i].colour[2] = s->aux_light_pre_colour[i].B;
      lights[i].enabled = 1;
    } else {
      lights[i].enabled = 0;
    }
  }
  return n + 4;
}

int oracle_scene_has_lights(const mbr_scene *s) {
  if (s->aux_light_number > 0) return 1;
  for (int i = 0; i < 4; i++)
    if (s->aux_light_pre_enabled[i]) return 1;
  return 0;
}

static void light_set_init(light_set *L, const mbr_scene *s, const mbr_light *tab) {
  L->tab = tab;
  L->number = s->aux_light_number > 0 ? s->aux_light_number : 0;
  L->count = L->number < 4 ? 4 : L->number;
  L->any = 0;
  for (int i = 0; i < L->count; i++)
    if (light_used(L, i)) L->any = 1;
}

/* FakeLights (src/shaders.cpp:1225-1270): the orbit-trap sum at the point
 * and one threshold-sized step along each axis gives the fake light's
 * direction; Lambert and a power-30 specular term, intensity / sum.
 * Parity partly pinned: the orbit trap by closed forms, this shader and the
 * glow only by linearity in intensity and visibility. */
static shade_t fake_lights(shader_in *in, shade_t *fakeSpec) {
  const mbr_scene *s = in->f->s;
  shade_t out = {0, 0, 0, 0};
  int it = 0;
  double delta = in->dist_thresh * s->smoothness;
  double rr = compute(in->f, MODE_ORBIT_TRAP, in->point, s->N, &it);
  double fakeLight = s->fake_lights_intensity / rr;
  double r = 1.0 / (rr + 1e-30);
  double rx = 1.0 / (compute(in->f, MODE_ORBIT_TRAP, v3add(in->point, v3m(delta, 0.0, 0.0)), s->N, &it) + 1e-30);
  double ry = 1.0 / (compute(in->f, MODE_ORBIT_TRAP, v3add(in->point, v3m(0.0, delta, 0.0)), s->N, &it) + 1e-30);
  double rz = 1.0 / (compute(in->f, MODE_ORBIT_TRAP, v3add(in->point, v3m(0.0, 0.0, delta)), s->N, &it) + 1e-30);
  v3 n = v3m(r - rx, r - ry, r - rz);
  if (n.x == 0 && n.y == 0 && n.z == 0) n.x = 0.0;
  else n = v3norm(n);
  double fakeLight2 = fakeLight * v3dot(in->normal, n);
  if (fakeLight2 < 0) fakeLight2 = 0;
  out.R = out.G = out.B = fakeLight2;
  v3 half = v3norm(v3sub(n, in->viewVector));
  double fs = v3dot(in->normal, half);
  if (fs < 0.0) fs = 0.0;
  fs = pow(fs, 30.0) * fakeLight;
  if (fs > 15.0) fs = 15.0;
  fakeSpec->R = fakeSpec->G = fakeSpec->B = fs;
  return out;
}

/* The textures of the textured background and environment mapping (set
 * by oracle_set_textures; NULL = none). */
static const mbr_texture *g_bg_tex = NULL, *g_env_tex = NULL;

void oracle_set_textures(const mbr_texture *bg, const mbr_texture *env) {
  g_bg_tex = bg;
  g_env_tex = env;
}

/* cTexture::Pixel / Interpolation (src/texture.cpp:83-111): bilinear over
 * four texels of the flat bitmap (ix + 1 at the last column is the next
 * row's first texel), truncated to 8 bits, black outside [0, w) x [0, h-1).
 * The texel one past the bitmap's end, which the reference reads at the
 * last column of the second-last row, reads black here. */
static void texture_pixel(const mbr_texture *t, double x, double y, int out[3]) {
  out[0] = out[1] = out[2] = 0;
  if (!(x >= 0 && x < t->width && y >= 0 && y < t->height - 1.0)) return;
  int ix = (int)x, iy = (int)y;
  double rx = (x - (int)x), ry = (y - (int)y);
  long long n = (long long)t->width * t->height;
  long long i1 = (long long)iy * t->width + ix, i2 = i1 + 1;
  long long i3 = (long long)(iy + 1) * t->width + ix, i4 = i3 + 1;
  for (int c = 0; c < 3; c++) {
    int k1 = t->rgb[3 * i1 + c], k2 = t->rgb[3 * i2 + c], k3 = t->rgb[3 * i3 + c];
    int k4 = i4 < n ? t->rgb[3 * i4 + c] : 0;
    out[c] = (unsigned char)(k1 * (1.0 - rx) * (1.0 - ry) + k2 * (rx) * (1.0 - ry) + k3 * (1.0 - rx) * ry + k4 * (rx * ry));
  }
}

/* EnvMapping (src/shaders.cpp:785-812): the view vector mirrored about the
 * normal indexes the environment texture by its angles, wrapped with the
 * reference's 8.25 and 8 texture offsets. */
static shade_t env_mapping(v3 normal, v3 viewVector) {
  shade_t out = {0, 0, 0, 0};
  if (!g_env_tex) return out;
  double dot = -v3dot(viewVector, normal);
  v3 reflect = v3add(v3mul(v3mul(normal, 2.0), dot), viewVector);
  double alfaTexture = atan2(reflect.y, reflect.x) + M_PI;
  double betaTexture = atan2(reflect.z, sqrt(reflect.x * reflect.x + reflect.y * reflect.y));
  double texWidth = g_env_tex->width, texHeight = g_env_tex->height;
  if (betaTexture > 0.5 * M_PI) betaTexture = 0.5 * M_PI - betaTexture;
  if (betaTexture < -0.5 * M_PI) betaTexture = -0.5 * M_PI + betaTexture;
  double dtx = (alfaTexture / (2.0 * M_PI)) * texWidth + texWidth * 8.25;
  double dty = (betaTexture / (M_PI) + 0.5) * texHeight + texHeight * 8.0;
  dtx = fmod(dtx, texWidth);
  dty = fmod(dty, texHeight);
  if (dtx < 0) dtx = 0;
  if (dty < 0) dty = 0;
  int p[3];
  texture_pixel(g_env_tex, dtx, dty, p);
  out.R = p[0] / 256.0;
  out.G = p[1] / 256.0;
  out.B = p[2] / 256.0;
  return out;
}

void oracle_env_mapping(const double normal[3], const double view[3], double rgb[3]) {
  shade_t e = env_mapping(v3m(normal[0], normal[1], normal[2]), v3m(view[0], view[1], view[2]));
  rgb[0] = e.R; rgb[1] = e.G; rgb[2] = e.B;
}

/* ObjectShader (src/shaders.cpp:21-112). */
static shade_t object_shader(shader_in *in, shade_t *surfaceColour, shade_t *specularOut) {
  const mbr_scene *s = in->f->s;
  shade_t output;
  v3 vn = calculate_normals(in);
  in->normal = vn;
  shade_t mainLight;
  mainLight.R = s->main_light_intensity * s->main_light_colour.R / 65536.0;
  mainLight.G = s->main_light_intensity * s->main_light_colour.G / 65536.0;
  mainLight.B = s->main_light_intensity * s->main_light_colour.B / 65536.0;
  /* MainShading (:759-769) */
  double sh = v3dot(in->normal, in->lightVect);
  if (sh < 0) sh = 0;
  double shade = s->direct_light * ((1.0 - s->shading) + s->shading * sh);
  double shadow = 1.0;
  if (s->shadow) shadow = main_shadow(in);
  /* MainSpecular (:771-784) */
  v3 half = v3norm(v3sub(in->lightVect, in->viewVector));
  double shade2 = v3dot(in->normal, half);
  if (shade2 < 0.0) shade2 = 0.0;
  shade2 = pow(shade2, 30.0) * 1.0;
  if (shade2 > 15.0) shade2 = 15.0;
  double specular = shade2 * s->specular;
  shade_t colour = surface_colour(in);
  *surfaceColour = colour;
  shade_t ambient = {0, 0, 0, 0};
  if (s->global_illumination) {
    if (s->fast_global_illumination) {
      double ao = fast_ao2(in);
      ambient.R = ambient.G = ambient.B = ao;
    } else {
      ambient = ambient_occlusion(in);
    }
  }
  shade_t ambient2;
  ambient2.R = ambient.R * s->global_ilum;
  ambient2.G = ambient.G * s->global_ilum;
  ambient2.B = ambient.B * s->global_ilum;
  shade_t envMapping = {0, 0, 0, 0};
  if (!s->raytraced_reflections && s->reflect > 0) envMapping = env_mapping(in->normal, in->viewVector);
  envMapping.R *= s->reflect;
  envMapping.G *= s->reflect;
  envMapping.B *= s->reflect;
  output.R = envMapping.R * ambient.R + (ambient2.R + mainLight.R * shade * shadow) * colour.R;
  output.G = envMapping.G * ambient.G + (ambient2.G + mainLight.G * shade * shadow) * colour.G;
  output.B = envMapping.B * ambient.B + (ambient2.B + mainLight.B * shade * shadow) * colour.B;
  shade_t auxSpecular;
  shade_t auxLights = aux_lights_shader(in, &auxSpecular);
  shade_t fakeLights = {0, 0, 0, 0}, fakeSpecular = {0, 0, 0, 0};
  if (s->fake_lights_enabled) fakeLights = fake_lights(in, &fakeSpecular);
  output.R += (auxLights.R + fakeLights.R) * colour.R;
  output.G += (auxLights.G + fakeLights.G) * colour.G;
  output.B += (auxLights.B + fakeLights.B) * colour.B;
  output.alpha = 1.0;
  specularOut->R = auxSpecular.R + fakeSpecular.R + mainLight.R * specular * shadow;
  specularOut->G = auxSpecular.G + fakeSpecular.G + mainLight.G * specular * shadow;
  specularOut->B = auxSpecular.B + fakeSpecular.B + mainLight.B * specular * shadow;
  specularOut->alpha = output.alpha;
  return output;
}

/* BackgroundShader (src/shaders.cpp:114-183): the texture branches (a
 * fisheye dome or an equirectangular map indexed by the view vector's
 * angles; the alpha the reference leaves unset is 0 here) and the gradient,
 * stored in sRGB16 before scaling. */
static shade_t background_shader(const mbr_scene *s, v3 viewVector) {
  shade_t px;
  if (s->textured_background && g_bg_tex) {
    double alfa = atan2(viewVector.y, viewVector.x);
    double beta = atan2(viewVector.z, sqrt(viewVector.x * viewVector.x + viewVector.y * viewVector.y));
    double texX, texY;
    if (s->background_as_fulldome) {
      int texWidth = g_bg_tex->width * 0.5;
      int texHeight = g_bg_tex->height;
      int offset = 0;
      if (beta < 0) {
        beta = -beta;
        alfa += M_PI;
        offset = texWidth;
      }
      texX = 0.5 * texWidth + cos(alfa) * (1.0 - beta / (0.5 * M_PI)) * texWidth * 0.5 + offset;
      texY = 0.5 * texHeight + sin(alfa) * (1.0 - beta / (0.5 * M_PI)) * texHeight * 0.5;
    } else {
      alfa = fmod(alfa + 2.5 * M_PI, 2 * M_PI);
      if (beta > 0.5 * M_PI) beta = 0.5 * M_PI - beta;
      if (beta < -0.5 * M_PI) beta = -0.5 * M_PI + beta;
      texX = alfa / (2.0 * M_PI) * g_bg_tex->width;
      texY = (beta / (M_PI) + 0.5) * g_bg_tex->height;
    }
    int p[3];
    texture_pixel(g_bg_tex, texX, texY, p);
    px.R = p[0] / 256.0;
    px.G = p[1] / 256.0;
    px.B = p[2] / 256.0;
    px.alpha = 0.0;
    return px;
  }
  v3 vec = v3norm(v3m(0.0, 0.0, -1.0));
  double grad = (v3dot(viewVector, vec) + 1.0);
  unsigned short R, G, B;
  if (grad < 1) {
    double Ngrad = 1.0 - grad;
    R = x86_d2u16(s->background_color3.R * Ngrad + s->background_color2.R * grad);
    G = x86_d2u16(s->background_color3.G * Ngrad + s->background_color2.G * grad);
    B = x86_d2u16(s->background_color3.B * Ngrad + s->background_color2.B * grad);
  } else {
    grad = grad - 1;
    double Ngrad = 1.0 - grad;
    R = x86_d2u16(s->background_color2.R * Ngrad + s->background_color1.R * grad);
    G = x86_d2u16(s->background_color2.G * Ngrad + s->background_color1.G * grad);
    B = x86_d2u16(s->background_color2.B * Ngrad + s->background_color1.B * grad);
  }
  px.R = R / 65536.0; px.G = G / 65536.0; px.B = B / 65536.0;
  px.alpha = 0.0;
  return px;
}

/* VolumetricShader (src/shaders.cpp:185-478): glow, basic fog and
 * volumetric fog (visible aux lights contribute nothing when no light is
 * enabled; volumetric lights and iteration fog are rejected
 * scene features).  *opacity_written is 0 when no step group was emitted
 * (the reference then leaves opacityOut uninitialised; we report 0). */
static shade_t volumetric_shader(const mbr_scene *s, double depth, const step_rec *buf, int stepCount,
                                 shade_t oldPixel, const light_set *L, v3 viewVector, const shader_in *pin,
                                 double *opacityOut) {
  shade_t output = oldPixel;
  /* the shader's own copy of the pixel's shader input (the reference passes
   * it by value); point and threshold follow the step records */
  shader_in in;
  if (pin) in = *pin;
  const int volLights = pin && (s->volumetric_light_enabled[0] || s->volumetric_light_enabled[1] ||
                                s->volumetric_light_enabled[2] || s->volumetric_light_enabled[3] ||
                                s->volumetric_light_enabled[4]);
  const int iterFog = pin && s->iter_fog_enabled;
  float totalOpacity = 0.0f;
  double fogVisibility = pow(10.0, s->fog_visibility / 10 - 16.0);
  double colourThresh = s->fog_colour1_distance;
  double colourThresh2 = s->fog_colour2_distance;
  double fogReduce = s->fog_distance_factor;
  double fogIntensity = s->fog_density;
  double glow = stepCount * s->glow_intensity / 512.0 * s->DE_factor;
  double glowN = 1.0 - glow;
  if (glowN < 0.0) glowN = 0.0;
  double glowR = (s->glow_color1.R * glowN + s->glow_color2.R * glow) / 65536.0;
  double glowG = (s->glow_color1.G * glowN + s->glow_color2.G * glow) / 65536.0;
  double glowB = (s->glow_color1.B * glowN + s->glow_color2.B * glow) / 65536.0;
  double totalStep = 0.0;
  *opacityOut = 0.0;
  for (int index = stepCount - 1; index > 0; index--) {
    double step = buf[index].step;
    double distance = buf[index].distance;
    totalStep += step;
    double dist_thresh = buf[index].distThresh;
    in.point = buf[index].point;
    in.dist_thresh = dist_thresh;
    if (totalStep < dist_thresh * s->quality) continue;
    step = totalStep;
    totalStep = 0.0;

    double glowOpacity = glow * step / depth;
    output.R = glowOpacity * glowR + (1.0 - glowOpacity) * output.R;
    output.G = glowOpacity * glowG + (1.0 - glowOpacity) * output.G;
    output.B = glowOpacity * glowB + (1.0 - glowOpacity) * output.B;
    output.alpha += glowOpacity;

    /* visible aux lights (:241-300); without a light taking part the block
     * changes nothing and is skipped */
    if (s->aux_light_visibility > 0 && L && L->any) {
      v3 point = buf[index].point;
      double lowestLightSize = 1e10;
      double lowestLightDist = 1e10;
      for (int i = 0; i < L->count; ++i) {
        if (!light_used(L, i)) continue;
        double lightDist = v3len(v3sub(point, v3from(L->tab[i].position)));
        double lightSize = L->tab[i].intensity * s->aux_light_intensity * s->aux_light_visibility;
        double r2 = lightDist / lightSize;
        if (r2 < lowestLightSize) {
          lowestLightSize = r2;
          lowestLightDist = lightDist;
        }
      }
      int smallSteps = x86_d2i(10.0 * step / (lowestLightDist + 1.0e-20));
      if (smallSteps > 50) smallSteps = 50;
      int smallStep_start = 0, smallStep_end = 1;
      double step2 = step;
      if (smallSteps > 0) {
        smallStep_start = 1;
        smallStep_end = smallSteps + 1;
        step2 = step / (smallSteps + 1.0);
      }
      for (int smallStep = smallStep_start; smallStep < smallStep_end; smallStep++) {
        v3 point2 = v3sub(point, v3mul(v3mul(viewVector, step2), (double)smallStep));
        for (int i = 0; i < L->count; ++i) {
          if (!light_used(L, i)) continue;
          double lightDist = v3len(v3sub(point2, v3from(L->tab[i].position)));
          double lightSize = L->tab[i].intensity * s->aux_light_intensity * s->aux_light_visibility;
          double r2 = lightDist / lightSize;
          if (r2 > 1.0) r2 = 1.0;
          double bellFunction = (cos(r2 * M_PI) + 1.0) / (r2 * r2 + 0.02) * 0.3;
          double lightDensity = step2 * bellFunction / lightSize;
          output.R += lightDensity * L->tab[i].colour[0] / 65536.0;
          output.G += lightDensity * L->tab[i].colour[1] / 65536.0;
          output.B += lightDensity * L->tab[i].colour[2] / 65536.0;
          output.alpha += lightDensity;
        }
      }
    }

    /* fake lights' glow (:302-312): the orbit-trap sum at the group's point
     * (1e-30f is the reference's float literal) */
    if (s->fake_lights_enabled && pin) {
      int it = 0;
      double rt = compute(in.f, MODE_ORBIT_TRAP, buf[index].point, s->N, &it);
      rt = sqrt(1.0f / (rt + 1.0e-30f));
      double fakeLight = 1.0 / (pow(rt, 10.0 / s->fake_lights_visibility_size) *
                                    pow(10.0, 10.0 / s->fake_lights_visibility_size) + 1e-100);
      output.R += fakeLight * step * s->fake_lights_visibility;
      output.G += fakeLight * step * s->fake_lights_visibility;
      output.B += fakeLight * step * s->fake_lights_visibility;
      output.alpha += fakeLight * step * s->fake_lights_visibility;
    }

    /* volumetric lights with shadows in fog (:314-341): the main light
     * (slot 0) and aux lights 1-4 (Lights[0..3], when enabled) */
    if (volLights) {
      for (int i = 0; i < 5; i++) {
        if (i == 0 && s->volumetric_light_enabled[0]) {
          double sh = main_shadow(&in);
          output.R += sh * step * s->volumetric_light_intensity[0] * s->main_light_colour.R / 65536.0;
          output.G += sh * step * s->volumetric_light_intensity[0] * s->main_light_colour.G / 65536.0;
          output.B += sh * step * s->volumetric_light_intensity[0] * s->main_light_colour.B / 65536.0;
          output.alpha += (sh + sh + sh) / 3.0 * step * s->volumetric_light_intensity[0];
        }
        if (i > 0 && L && L->tab[i - 1].enabled && s->volumetric_light_enabled[i]) {
          const mbr_light *lt = &L->tab[i - 1];
          v3 d = v3sub(v3from(lt->position), in.point);
          double dl = v3len(d);
          double distance2 = dl * dl;
          v3 lightVectorTemp = v3norm(d);
          double light = aux_shadow(&in, dl, lightVectorTemp);
          output.R += light * lt->colour[0] / 65536.0 * s->volumetric_light_intensity[i] * step / distance2;
          output.G += light * lt->colour[1] / 65536.0 * s->volumetric_light_intensity[i] * step / distance2;
          output.B += light * lt->colour[2] / 65536.0 * s->volumetric_light_intensity[i] * step / distance2;
          output.alpha += light * s->volumetric_light_intensity[i] * step / distance2;
        }
      }
    }

    if (s->post_fog_enabled) {
      double fogDensity = step / fogVisibility;
      if (fogDensity > 1.0) fogDensity = 1.0;
      output.R = fogDensity * s->post_fog_color.R / 65536.0 + (1.0 - fogDensity) * output.R;
      output.G = fogDensity * s->post_fog_color.G / 65536.0 + (1.0 - fogDensity) * output.G;
      output.B = fogDensity * s->post_fog_color.B / 65536.0 + (1.0 - fogDensity) * output.B;
      totalOpacity = fogDensity + (1.0 - fogDensity) * totalOpacity;
      output.alpha = fogDensity + (1.0 - fogDensity) * output.alpha;
    }

    if (fogIntensity > 0.0) {
      double densityTemp = (step * fogReduce) / (distance * distance + fogReduce * fogReduce);
      double k = distance / colourThresh;
      if (k > 1) k = 1.0;
      double kn = 1.0 - k;
      double fogRtemp = (s->fog_colour1.R * kn + s->fog_colour2.R * k);
      double fogGtemp = (s->fog_colour1.G * kn + s->fog_colour2.G * k);
      double fogBtemp = (s->fog_colour1.B * kn + s->fog_colour2.B * k);
      double k2 = distance / colourThresh2 * k;
      if (k2 > 1) k2 = 1.0;
      kn = 1.0 - k2;
      fogRtemp = (fogRtemp * kn + s->fog_colour3.R * k2);
      fogGtemp = (fogGtemp * kn + s->fog_colour3.G * k2);
      fogBtemp = (fogBtemp * kn + s->fog_colour3.B * k2);
      double fogDensity = 0.3 * fogIntensity * densityTemp / (1.0 + fogIntensity * densityTemp);
      if (fogDensity > 1) fogDensity = 1.0;
      output.R = fogDensity * fogRtemp / 65536.0 + (1.0 - fogDensity) * output.R;
      output.G = fogDensity * fogGtemp / 65536.0 + (1.0 - fogDensity) * output.G;
      output.B = fogDensity * fogBtemp / 65536.0 + (1.0 - fogDensity) * output.B;
      totalOpacity = fogDensity + (1.0 - fogDensity) * totalOpacity;
      output.alpha = fogDensity + (1.0 - fogDensity) * output.alpha;
    }

    /* iteration fog (:385-465) */
    if (iterFog) {
      int Lit = buf[index].iters;
      double opacity = iter_opacity(step, Lit, s->N, s->iter_fog_opacity_trim, s->iter_fog_opacity);
      shade_t newColour = {0.0, 0.0, 0.0, 0.0};
      if (opacity > 0) {
        double iterFactor = (double)2.0 * (Lit - s->iter_fog_opacity_trim) / (s->N - s->iter_fog_opacity_trim);
        double k = iterFactor;
        if (k > 1.0) k = 1.0;
        if (k < 0.0) k = 0.0;
        double kn = 1.0 - k;
        double fogColR = (s->fog_colour1.R * kn + s->fog_colour2.R * k);
        double fogColG = (s->fog_colour1.G * kn + s->fog_colour2.G * k);
        double fogColB = (s->fog_colour1.B * kn + s->fog_colour2.B * k);
        double k2 = iterFactor - 1.0;
        if (k2 < 0.0) k2 = 0.0;
        if (k2 > 1.0) k2 = 1.0;
        kn = 1.0 - k2;
        fogColR = (fogColR * kn + s->fog_colour3.R * k2);
        fogColG = (fogColG * kn + s->fog_colour3.G * k2);
        fogColB = (fogColB * kn + s->fog_colour3.B * k2);
        for (int i = 0; i < 5; i++) {
          if (i == 0) {
            if (s->main_light_intensity * s->direct_light > 0.0) {
              double sh = main_shadow(&in);
              newColour.R += sh * s->main_light_colour.R / 65536.0 * s->main_light_intensity * s->direct_light;
              newColour.G += sh * s->main_light_colour.G / 65536.0 * s->main_light_intensity * s->direct_light;
              newColour.B += sh * s->main_light_colour.B / 65536.0 * s->main_light_intensity * s->direct_light;
            }
          }
          if (i > 0 && L && L->tab[i - 1].enabled) {
            const mbr_light *lt = &L->tab[i - 1];
            v3 d = v3sub(v3from(lt->position), in.point);
            double dl = v3len(d);
            double distance2 = dl * dl;
            v3 lightVectorTemp = v3norm(d);
            double light = aux_shadow(&in, dl, lightVectorTemp);
            double intensity = lt->intensity * 100.0;
            newColour.R += light * lt->colour[0] / 65536.0 / distance2 * intensity;
            newColour.G += light * lt->colour[1] / 65536.0 / distance2 * intensity;
            newColour.B += light * lt->colour[2] / 65536.0 / distance2 * intensity;
          }
        }
        if (s->global_illumination && !s->fast_global_illumination) {
          shade_t AO = ambient_occlusion(&in);
          newColour.R += AO.R * s->global_ilum;
          newColour.G += AO.G * s->global_ilum;
          newColour.B += AO.B * s->global_ilum;
        }
        if (opacity > 1.0) opacity = 1.0;
        output.R = output.R * (1.0 - opacity) + newColour.R * opacity * fogColR / 65536.0;
        output.G = output.G * (1.0 - opacity) + newColour.G * opacity * fogColG / 65536.0;
        output.B = output.B * (1.0 - opacity) + newColour.B * opacity * fogColB / 65536.0;
        totalOpacity = opacity + (1.0 - opacity) * totalOpacity;
        output.alpha = opacity + (1.0 - opacity) * output.alpha;
      }
    }

    if (totalOpacity > 1.0) totalOpacity = 1.0;
    if (output.alpha > 1.0) output.alpha = 1.0;
    *opacityOut = totalOpacity;
  }
  return output;
}

/* ------------------------------------------------------------ AO table */
int oracle_ao_table(int AO_quality, const mbr_texture *lm, double *dirs, int *rgb, int max) {
  /* src/Render3D.cpp:271-313; white 100x100 texture when lm == NULL */
  int counter = 0;
  int W = lm ? lm->width : 100, H = lm ? lm->height : 100;
  for (double b = -0.49 * M_PI; b < 0.49 * M_PI; b += 1.0 / AO_quality) {
    for (double a = 0.0; a < 2.0 * M_PI; a += ((2.0 / AO_quality) / cos(b))) {
      double d[3] = {cos(a + b) * cos(b), sin(a + b) * cos(b), sin(b)};
      int X = (int)((a + b) / (2.0 * M_PI) * W + W * 8.5) % W;
      int Y = (int)(b / (M_PI) * H + H * 8.5) % H;
      int R = 255, G = 255, B = 255;
      if (lm) {
        const uint8_t *px = lm->rgb + 3 * ((size_t)X + (size_t)Y * W);
        R = px[0]; G = px[1]; B = px[2];
      }
      if (counter < max) {
        dirs[3 * counter + 0] = d[0]; dirs[3 * counter + 1] = d[1]; dirs[3 * counter + 2] = d[2];
        rgb[3 * counter + 0] = R; rgb[3 * counter + 1] = G; rgb[3 * counter + 2] = B;
      }
      if (R > 10 || G > 10 || B > 10) counter++;
      if (counter >= 10000) break;
    }
    if (counter >= 10000) break;
  }
  if (counter == 0) {
    counter = 1;
    dirs[0] = dirs[1] = dirs[2] = 0.0;
    rgb[0] = rgb[1] = rgb[2] = 0;
  }
  return counter;
}

/* --------------------------------------------------------- public units */
double oracle_compute(const mbr_scene *s, int mode, const double p[3], int n_iter, int *iters) {
  fractal_ctx f; fractal_ctx_init(&f, s);
  return compute(&f, mode, v3from(p), n_iter, iters);
}

double oracle_calculate_distance(const mbr_scene *s, const double p[3], double detail_size,
                                 int n_iter, int min_n, int normal_mode,
                                 int *max_iter, int *iters, int *object) {
  fractal_ctx f; fractal_ctx_init(&f, s);
  int mi = 0, it = 0, ob = 0;
  double d = calc_distance(&f, v3from(p), detail_size, n_iter, min_n, normal_mode, &mi, &it, &ob, NULL);
  if (max_iter) *max_iter = mi;
  if (iters) *iters = it;
  if (object) *object = ob;
  return d;
}

void oracle_camera(const mbr_scene *s, double R[9], double start[3], double light[3]) {
  camera_t c; camera_init(&c, s);
  v3 ex = rot_apply(&c.R, v3m(1, 0, 0)), ey = rot_apply(&c.R, v3m(0, 1, 0)), ez = rot_apply(&c.R, v3m(0, 0, 1));
  /* columns are the images of the base vectors */
  R[0] = ex.x; R[1] = ey.x; R[2] = ez.x;
  R[3] = ex.y; R[4] = ey.y; R[5] = ez.y;
  R[6] = ex.z; R[7] = ey.z; R[8] = ez.z;
  start[0] = c.start.x; start[1] = c.start.y; start[2] = c.start.z;
  light[0] = c.light.x; light[1] = c.light.y; light[2] = c.light.z;
}

int oracle_view_vector(const mbr_scene *s, int x, int y, double dir[3]) {
  camera_t c; camera_init(&c, s);
  v3 v; int cut = view_vector(s, &c, x, y, &v);
  dir[0] = v.x; dir[1] = v.y; dir[2] = v.z;
  return cut;
}

int oracle_march(const mbr_scene *s, const double start[3], const double dir[3], uint32_t pixel,
                 oracle_march_out *out) {
  fractal_ctx f; fractal_ctx_init(&f, s);
  camera_t c; camera_init(&c, s);
  step_rec *buf = (step_rec *)malloc(sizeof(step_rec) * 10002);
  if (!buf) return -1;
  jitter_t j = {s->jitter_mode == MBR_JITTER_HASH ? ORACLE_JITTER_HASH : ORACLE_JITTER_NONE,
                s->jitter_seed, (uint32_t)s->frame_no, pixel};
  march_result m;
  ray_march(&f, &c, v3from(start), v3from(dir), s->view_distance_max, 0.0, &j, buf, &m, NULL);
  out->point[0] = m.point.x; out->point[1] = m.point.y; out->point[2] = m.point.z;
  out->depth = m.depth; out->dist_thresh = m.distThresh; out->last_dist = m.lastDist;
  out->found = m.found; out->step_count = m.buffCount; out->object = m.object; out->de_calls = m.de_calls;
  free(buf);
  return 0;
}

void oracle_rotation3(const double angles[3], double m[9]) {
  rotm r;
  rot_set3(&r, angles);
  memcpy(m, r.m, sizeof r.m);
}

void oracle_volumetric(const mbr_scene *s, double depth, const double *steps, int n,
                       const double pixel[4], double out[4], double *opacity) {
  step_rec *buf = (step_rec *)calloc(n > 0 ? n : 1, sizeof(step_rec));
  for (int i = 0; i < n; i++) {
    buf[i].step = steps[3 * i + 0];
    buf[i].distance = steps[3 * i + 1];
    buf[i].distThresh = steps[3 * i + 2];
  }
  shade_t px = {pixel[0], pixel[1], pixel[2], pixel[3]};
  shade_t o = volumetric_shader(s, depth, buf, n, px, NULL, v3m(0, 0, 0), NULL, opacity);
  out[0] = o.R; out[1] = o.G; out[2] = o.B; out[3] = o.alpha;
  free(buf);
}

void oracle_index_to_colour(const mbr_scene *s, int index, int rgb[3]) {
  mbr_rgb c = index_to_colour(index, s->palette);
  rgb[0] = c.R; rgb[1] = c.G; rgb[2] = c.B;
}

void oracle_background(const mbr_scene *s, const double view[3], double rgb[3]) {
  shade_t b = background_shader(s, v3from(view));
  rgb[0] = b.R; rgb[1] = b.G; rgb[2] = b.B;
}

int oracle_object_shader(const mbr_scene *s, const double point[3], const double view[3],
                         double dist_thresh, double depth, double last_dist, int object,
                         oracle_shade_out *out) {
  fractal_ctx f; fractal_ctx_init(&f, s);
  camera_t c; camera_init(&c, s);
  double *dirs = (double *)malloc(sizeof(double) * 3 * 10000);
  int *rgb = (int *)malloc(sizeof(int) * 3 * 10000);
  int q = s->ao_quality; if (q >= 30) q = 30; if (q <= 1) q = 1;
  int count = oracle_ao_table(q, NULL, dirs, rgb, 10000);
  shader_in in;
  memset(&in, 0, sizeof in);
  in.f = &f; in.c = &c;
  in.point = v3from(point);
  in.viewVector = v3from(view);
  in.lightVect = c.light;
  in.dist_thresh = dist_thresh;
  in.delta = s->constant_DE_threshold ? depth * c.resolution * s->persp : dist_thresh * s->quality;
  in.detail = dist_thresh;
  in.lastDist = last_dist;
  in.depth = depth;
  in.objectType = object;
  in.ao_dirs = dirs; in.ao_rgb = rgb; in.ao_count = count;
  mbr_light *tab = NULL;
  light_set L;
  if (oracle_scene_has_lights(s)) {
    int n = (s->aux_light_number > 0 ? s->aux_light_number : 0) + 4;
    tab = (mbr_light *)calloc((size_t)n, sizeof(mbr_light));
    if (!tab || oracle_place_lights(s, tab, n) != n) {
      free(tab); free(dirs); free(rgb);
      return -1;
    }
    light_set_init(&L, s, tab);
    in.lights = &L;
  }
  shade_t colour, spec;
  shade_t o = object_shader(&in, &colour, &spec);
  out->normal[0] = in.normal.x; out->normal[1] = in.normal.y; out->normal[2] = in.normal.z;
  /* recompute the scalar terms for inspection (same functions, same state) */
  double sh = v3dot(in.normal, in.lightVect);
  if (sh < 0) sh = 0;
  out->shade = s->direct_light * ((1.0 - s->shading) + s->shading * sh);
  in.detail = dist_thresh;
  out->shadow = s->shadow ? main_shadow(&in) : 1.0;
  v3 half = v3norm(v3sub(in.lightVect, in.viewVector));
  double s2 = v3dot(in.normal, half);
  if (s2 < 0) s2 = 0;
  s2 = pow(s2, 30.0);
  if (s2 > 15.0) s2 = 15.0;
  out->specular = s2 * s->specular;
  shade_t ao = {0, 0, 0, 0};
  if (s->global_illumination) {
    if (s->fast_global_illumination) { double a = fast_ao2(&in); ao.R = ao.G = ao.B = a; }
    else ao = ambient_occlusion(&in);
  }
  out->ao[0] = ao.R; out->ao[1] = ao.G; out->ao[2] = ao.B;
  out->colour[0] = colour.R; out->colour[1] = colour.G; out->colour[2] = colour.B;
  out->rgb[0] = o.R; out->rgb[1] = o.G; out->rgb[2] = o.B;
  out->spec_rgb[0] = spec.R; out->spec_rgb[1] = spec.G; out->spec_rgb[2] = spec.B;
  free(dirs); free(rgb); free(tab);
  return 0;
}

/* ---------------------------------------------------------------- render */
typedef struct {
  const mbr_scene *s;
  const fractal_ctx *f;
  const camera_t *c;
  const double *ao_dirs; const int *ao_rgb; int ao_count;
  mbr_framebuffer *fb;
  int row_begin, row_end;
  int jitter; uint32_t glibc_seed;
  const light_set *lights;
  volatile int next_row;
  pthread_mutex_t lock;
  oracle_stats st;
} render_job;

/* per-pixel body of MainThread (src/Render3D.cpp:387-590) for ray 0 */
/* reflection coefficient of the surface a ray ended on (src/Render3D.cpp:471-477) */
static double object_reflect(const mbr_scene *s, int object) {
  switch (object) {
    case MBR_OBJ_FRACTAL: return s->reflect;
    case MBR_OBJ_WATER: return s->water_reflect;
    case MBR_OBJ_PLANE: return s->plane_reflect;
    case MBR_OBJ_BOX: return s->box_reflect;
    case MBR_OBJ_BOX_INV: return s->inv_box_reflect;
    case MBR_OBJ_SPHERE: return s->sphere_reflect;
    case MBR_OBJ_SPHERE_INV: return s->inv_sphere_reflect;
  }
  return 0.0;
}

typedef struct {
  v3 start, viewVector, point;
  double distThresh, lastDist, depth, reflect;
  int found, buffCount, objectType;
  step_rec *buf;
} refl_ray;

/* MainThread's pixel with raytraced reflections (src/Render3D.cpp:440-589):
 * march the primary ray and up to reflections_max mirrored rays (each from
 * its hit point along the view vector mirrored about the normal, with the
 * previous threshold as its initial one), then shade them from the last to
 * the first, each blending the shaded result of the next by its surface's
 * reflection coefficient, each through its own VolumetricShader. */
static void render_pixel_reflections(render_job *job, int x, int y, v3 viewVector, int cut, const jitter_t *j,
                                     step_rec *buf, oracle_stats *st, de_counter *mc, de_counter *sc) {
  const mbr_scene *s = job->s;
  const camera_t *c = job->c;
  const int W = s->image_width;
  const int maxRay = s->reflections_max;
  refl_ray rb[MBR_MAX_REFLECTIONS + 1];
  double distThresh = 0.0;
  v3 startRay = c->start;
  v3 point = v3m(0, 0, 0);
  int rayEnd = 0;
  for (int ray = 0; ray <= maxRay; ray++) {
    refl_ray *R = &rb[ray];
    R->buf = buf + (size_t)ray * 10002;
    R->start = startRay;
    R->viewVector = viewVector;
    R->found = 0;
    R->buffCount = 0;
    R->lastDist = 0;
    R->depth = 0;
    march_result m;
    memset(&m, 0, sizeof m);
    if (!cut) {
      ray_march(job->f, c, startRay, viewVector, s->view_distance_max, distThresh, j, R->buf, &m, mc);
      point = m.point;
      distThresh = m.distThresh;
    }
    st->rays++;
    if (m.found) st->hits++;
    st->missed_de += m.missed;
    if ((uint64_t)m.buffCount > st->max_steps) st->max_steps = m.buffCount;
    st->total_steps += m.buffCount;
    R->found = m.found;
    R->buffCount = m.buffCount;
    R->lastDist = m.lastDist;
    R->depth = m.depth;
    R->point = point;
    R->distThresh = distThresh;
    R->objectType = m.object;
    R->reflect = object_reflect(s, m.object);
    rayEnd = ray;
    if (!R->found) break;
    if (R->reflect == 0) break;
    /* the mirrored ray */
    shader_in in;
    memset(&in, 0, sizeof in);
    in.f = job->f; in.c = c;
    in.dist_thresh = R->distThresh;
    in.detail = R->distThresh;
    in.lightVect = c->light;
    in.point = point;
    in.viewVector = viewVector;
    in.delta = s->constant_DE_threshold ? R->depth * c->resolution * s->persp : R->distThresh * s->quality;
    in.cnt = sc;
    v3 vn = calculate_normals(&in);
    viewVector = v3sub(viewVector, v3mul(v3mul(vn, v3dot(viewVector, vn)), 2.0));
    startRay = v3add(point, v3mul(viewVector, R->distThresh));
  }

  shade_t resultShader = {0, 0, 0, 0};
  shade_t objectColour = {0, 0, 0, 0};
  double opacity0 = 0.0, alpha0 = 0.0;
  for (int ray = rayEnd; ray >= 0; ray--) {
    refl_ray *R = &rb[ray];
    shade_t objectShader = {0, 0, 0, 0}, backgroundShader = {0, 0, 0, 0}, specular = {0, 0, 0, 0};
    shader_in in;
    memset(&in, 0, sizeof in);
    in.f = job->f; in.c = c;
    in.dist_thresh = R->distThresh;
    in.delta = s->constant_DE_threshold ? R->depth * c->resolution * s->persp : R->distThresh * s->quality;
    in.detail = R->distThresh;
    in.lightVect = c->light;
    in.point = R->point;
    in.viewVector = R->viewVector;
    in.ao_dirs = job->ao_dirs; in.ao_rgb = job->ao_rgb; in.ao_count = job->ao_count;
    in.lights = job->lights;
    in.lastDist = R->lastDist;
    in.depth = R->depth;
    in.objectType = R->objectType;
    in.cnt = sc;
    const double depth = R->depth;
    if (R->found) objectShader = object_shader(&in, &objectColour, &specular);
    else {
      backgroundShader = background_shader(s, R->viewVector);
      R->depth = 1e20;
    }
    const double reflect = R->reflect;
    shade_t pixel;
    if (maxRay > 0 && rayEnd > 0 && ray != rayEnd) {
      pixel.R = resultShader.R * reflect + (1.0 - reflect) * (objectShader.R + backgroundShader.R) + specular.R;
      pixel.G = resultShader.G * reflect + (1.0 - reflect) * (objectShader.G + backgroundShader.G) + specular.G;
      pixel.B = resultShader.B * reflect + (1.0 - reflect) * (objectShader.B + backgroundShader.B) + specular.B;
    } else {
      pixel.R = objectShader.R + backgroundShader.R + specular.R;
      pixel.G = objectShader.G + backgroundShader.G + specular.G;
      pixel.B = objectShader.B + backgroundShader.B + specular.B;
    }
    pixel.alpha = objectShader.alpha + backgroundShader.alpha;
    double opacity;
    shade_t vol = volumetric_shader(s, depth, R->buf, R->buffCount, pixel, job->lights, R->viewVector, &in, &opacity);
    resultShader.R = vol.R;
    resultShader.G = vol.G;
    resultShader.B = vol.B;
    if (ray == 0) {
      opacity0 = opacity;
      alpha0 = vol.alpha;
    }
  }
  size_t idx = (size_t)x + (size_t)y * W;
  mbr_framebuffer *fb = job->fb;
  if (fb->rgb[0]) fb->rgb[0][idx] = (float)resultShader.R;
  if (fb->rgb[1]) fb->rgb[1][idx] = (float)resultShader.G;
  if (fb->rgb[2]) fb->rgb[2][idx] = (float)resultShader.B;
  if (fb->z) fb->z[idx] = (float)rb[0].depth;
  if (fb->colour[0]) fb->colour[0][idx] = x86_d2u8(objectColour.R * 255);
  if (fb->colour[1]) fb->colour[1][idx] = x86_d2u8(objectColour.G * 255);
  if (fb->colour[2]) fb->colour[2][idx] = x86_d2u8(objectColour.B * 255);
  if (fb->opacity) fb->opacity[idx] = x86_d2u16(opacity0 * 65535.0);
  if (fb->alpha) fb->alpha[idx] = x86_d2u16(alpha0 * 65535.0);
}

static void render_pixel(render_job *job, int x, int y, step_rec *buf, oracle_stats *st, de_counter *mc, de_counter *sc) {
  const mbr_scene *s = job->s;
  const camera_t *c = job->c;
  int W = s->image_width;
  v3 viewVector;
  int cut = view_vector(s, c, x, y, &viewVector);
  uint32_t pixel = (uint32_t)(x + y * W + s->tile_index * W * s->image_height);
  jitter_t j = {job->jitter, s->jitter_seed, (uint32_t)s->frame_no, pixel};
  march_result m;
  memset(&m, 0, sizeof m);
  if (s->raytraced_reflections && s->reflections_max > 0) {
    render_pixel_reflections(job, x, y, viewVector, cut, &j, buf, st, mc, sc);
    return;
  }
  if (!cut) ray_march(job->f, c, c->start, viewVector, s->view_distance_max, 0.0, &j, buf, &m, mc);
  st->rays++;
  if (m.found) st->hits++;
  st->missed_de += m.missed;
  if ((uint64_t)m.buffCount > st->max_steps) st->max_steps = m.buffCount;
  st->total_steps += m.buffCount;

  shade_t objectShader = {0, 0, 0, 0}, backgroundShader = {0, 0, 0, 0}, specular = {0, 0, 0, 0};
  shade_t objectColour = {0, 0, 0, 0};
  double zdepth = m.depth;
  shader_in in;
  memset(&in, 0, sizeof in);
  in.f = job->f; in.c = c;
  in.dist_thresh = m.distThresh;
  in.delta = s->constant_DE_threshold ? m.depth * c->resolution * s->persp : m.distThresh * s->quality;
  in.detail = m.distThresh;
  in.lightVect = c->light;
  in.point = m.point;
  in.viewVector = viewVector;
  in.ao_dirs = job->ao_dirs; in.ao_rgb = job->ao_rgb; in.ao_count = job->ao_count;
  in.lights = job->lights;
  in.lastDist = m.lastDist;
  in.depth = m.depth;
  in.objectType = m.object;
  in.cnt = sc;
  if (m.found) {
    objectShader = object_shader(&in, &objectColour, &specular);
  } else {
    backgroundShader = background_shader(s, viewVector);
    zdepth = 1e20;
  }
  shade_t pixel_sh;
  pixel_sh.R = objectShader.R + backgroundShader.R + specular.R;
  pixel_sh.G = objectShader.G + backgroundShader.G + specular.G;
  pixel_sh.B = objectShader.B + backgroundShader.B + specular.B;
  pixel_sh.alpha = objectShader.alpha + backgroundShader.alpha;
  double opacity;
  shade_t vol = volumetric_shader(s, m.depth, buf, m.buffCount, pixel_sh, job->lights, viewVector, &in, &opacity);

  size_t idx = (size_t)x + (size_t)y * W;
  mbr_framebuffer *fb = job->fb;
  if (fb->rgb[0]) fb->rgb[0][idx] = (float)vol.R;
  if (fb->rgb[1]) fb->rgb[1][idx] = (float)vol.G;
  if (fb->rgb[2]) fb->rgb[2][idx] = (float)vol.B;
  if (fb->z) fb->z[idx] = (float)zdepth;
  if (fb->colour[0]) fb->colour[0][idx] = x86_d2u8(objectColour.R * 255);
  if (fb->colour[1]) fb->colour[1][idx] = x86_d2u8(objectColour.G * 255);
  if (fb->colour[2]) fb->colour[2][idx] = x86_d2u8(objectColour.B * 255);
  if (fb->opacity) fb->opacity[idx] = x86_d2u16(opacity * 65535.0);
  if (fb->alpha) fb->alpha[idx] = x86_d2u16(vol.alpha * 65535.0);
}

static void *render_worker(void *arg) {
  render_job *job = (render_job *)arg;
  const int levels = (job->s->raytraced_reflections && job->s->reflections_max > 0) ? job->s->reflections_max + 1 : 1;
  step_rec *buf = (step_rec *)malloc(sizeof(step_rec) * 10002 * (size_t)levels);
  oracle_stats st; memset(&st, 0, sizeof st);
  de_counter mc = {0, 0}, sc = {0, 0};
  for (;;) {
    int y = __sync_fetch_and_add(&job->next_row, 1);
    if (y >= job->row_end) break;
    for (int x = 0; x < job->s->image_width; x++) render_pixel(job, x, y, buf, &st, &mc, &sc);
  }
  free(buf);
  st.march_de_calls = mc.calls; st.shade_de_calls = sc.calls; st.iterations = mc.iterations + sc.iterations;
  pthread_mutex_lock(&job->lock);
  job->st.rays += st.rays; job->st.hits += st.hits; job->st.missed_de += st.missed_de;
  job->st.march_de_calls += st.march_de_calls; job->st.shade_de_calls += st.shade_de_calls;
  job->st.iterations += st.iterations; job->st.total_steps += st.total_steps;
  if (st.max_steps > job->st.max_steps) job->st.max_steps = st.max_steps;
  pthread_mutex_unlock(&job->lock);
  return NULL;
}

int oracle_render(const mbr_scene *s, const mbr_texture *lightmap, int row_begin, int row_end,
                  int nthreads, int jitter, uint32_t glibc_seed, mbr_framebuffer *fb, oracle_stats *st) {
  if (!s || !fb || s->image_width <= 0 || s->image_height <= 0 || s->tiles < 1) return MBR_ERR_INVALID_ARG;
  if (row_end <= 0 || row_end > s->image_height) row_end = s->image_height;
  if (row_begin < 0) row_begin = 0;
  fractal_ctx f; fractal_ctx_init(&f, s);
  camera_t c; camera_init(&c, s);
  double *dirs = (double *)malloc(sizeof(double) * 3 * 10000);
  int *rgb = (int *)malloc(sizeof(int) * 3 * 10000);
  int q = s->ao_quality; if (q >= 30) q = 30; if (q <= 1) q = 1;
  int count = oracle_ao_table(q, lightmap, dirs, rgb, 10000);
  render_job job;
  memset(&job, 0, sizeof job);
  job.s = s; job.f = &f; job.c = &c; job.ao_dirs = dirs; job.ao_rgb = rgb; job.ao_count = count;
  job.fb = fb; job.row_begin = row_begin; job.row_end = row_end; job.next_row = row_begin;
  job.jitter = jitter; job.glibc_seed = glibc_seed;
  /* the lights are placed before the frame, as MainRender does
   * (src/Render3D.cpp:1990-1993) */
  mbr_light *tab = NULL;
  light_set L;
  memset(&L, 0, sizeof L);
  if (oracle_scene_has_lights(s)) {
    int n = (s->aux_light_number > 0 ? s->aux_light_number : 0) + 4;
    tab = (mbr_light *)calloc((size_t)n, sizeof(mbr_light));
    if (!tab || oracle_place_lights(s, tab, n) != n) {
      free(tab); free(dirs); free(rgb);
      return MBR_ERR_INVALID_ARG;
    }
    light_set_init(&L, s, tab);
    job.lights = &L;
  }
  pthread_mutex_init(&job.lock, NULL);
  if (jitter == ORACLE_JITTER_GLIBC) { nthreads = 1; srand(glibc_seed); }
  if (nthreads <= 0) { long n = sysconf(_SC_NPROCESSORS_ONLN); nthreads = n > 0 ? (int)n : 1; }
  if (nthreads == 1) {
    render_worker(&job);
  } else {
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
    for (int i = 0; i < nthreads; i++) pthread_create(&th[i], NULL, render_worker, &job);
    for (int i = 0; i < nthreads; i++) pthread_join(th[i], NULL);
    free(th);
  }
  pthread_mutex_destroy(&job.lock);
  if (st) *st = job.st;
  free(dirs); free(rgb); free(tab);
  return MBR_OK;
}

/* ---------------------------------------------------------------- SSAO */

/* Screen position of pixel (x, y) at depth z in the plane units of the
 * SSAO horizon test, for the three projections (src/image.cpp:288-314 for
 * the centre pixel, :337-362 for the samples).  The sample form keeps its
 * own operation order (z2 * persp grouped; fish-eye tests the loop radius,
 * not its own radius, before dividing). */
static void ssao_centre(int type, double fov, double aspect, int width, int height, int x, int y, double z,
                        double *px, double *py) {
  double x2, y2;
  if (type == MBR_PERSP_FISH_EYE) {
    x2 = M_PI * ((double)x / width - 0.5) * aspect;
    y2 = M_PI * ((double)y / height - 0.5);
    double r = sqrt(x2 * x2 + y2 * y2);
    if (r != 0.0) {
      x2 = x2 / r * sin(r * fov) * z;
      y2 = y2 / r * sin(r * fov) * z;
    }
  } else if (type == MBR_PERSP_EQUIRECTANGULAR) {
    x2 = M_PI * ((double)x / width - 0.5) * aspect;
    y2 = M_PI * ((double)y / height - 0.5);
    x2 = sin(fov * x2) * cos(fov * y2) * z;
    y2 = sin(fov * y2) * z;
  } else {
    x2 = ((double)x / width - 0.5) * aspect;
    y2 = ((double)y / height - 0.5);
    x2 = x2 * z * fov;
    y2 = y2 * z * fov;
  }
  *px = x2;
  *py = y2;
}

static void ssao_sample(int type, double fov, double aspect, int width, int height, double xx, double yy,
                        double z2, double *px, double *py) {
  double xx2, yy2;
  if (type == MBR_PERSP_FISH_EYE) {
    xx2 = M_PI * ((double)xx / width - 0.5) * aspect;
    yy2 = M_PI * ((double)yy / height - 0.5);
    double r2 = sqrt(xx2 * xx2 + yy2 * yy2);
    xx2 = xx2 / r2 * sin(r2 * fov) * z2;  /* loop radius >= 1: the guard always passes */
    yy2 = yy2 / r2 * sin(r2 * fov) * z2;
  } else if (type == MBR_PERSP_EQUIRECTANGULAR) {
    xx2 = M_PI * (xx / width - 0.5) * aspect;
    yy2 = M_PI * (yy / height - 0.5);
    xx2 = sin(fov * xx2) * cos(fov * yy2) * z2;
    yy2 = sin(fov * yy2) * z2;
  } else {
    xx2 = (xx / width - 0.5) * aspect;
    yy2 = (yy / height - 0.5);
    xx2 = xx2 * (z2 * fov);
    yy2 = yy2 * (z2 * fov);
  }
  *px = xx2;
  *py = yy2;
}

int oracle_ssao(const mbr_scene *s, int width, int height, const float *zb, const uint16_t *opacity,
                const uint8_t *const colour[3], uint16_t *image16) {
  const int q = s->ssao_quality;
  if (q < 1 || width <= 0 || height <= 0) return MBR_ERR_INVALID_ARG;
  double *sn = (double *)malloc(sizeof(double) * q), *cs = (double *)malloc(sizeof(double) * q);
  for (int i = 0; i < q; i++) {  /* image.cpp:266-271 */
    sn[i] = sin((double)i / q * 2.0 * M_PI);
    cs[i] = cos((double)i / q * 2.0 * M_PI);
  }
  const double scale = (double)width / (q * q) / 2.0;  /* image.cpp:273 */
  const double aspect = (double)width / height;
  const double fov = s->persp;
  const double intensity = s->global_ilum;
  const int type = s->perspective_type;
  for (int y = 0; y < height; y++) {
    for (int x = 0; x < width; x++) {
      const size_t p = (size_t)y * width + x;
      const double z = zb[p];
      const double op = opacity[p] / 65535.0;
      double total = 0;
      if (z < 1e19) {
        double x2, y2;
        ssao_centre(type, fov, aspect, width, height, x, y, z, &x2, &y2);
        double ambient = 0;
        for (int a = 0; a < q; a++) {
          double best = -1e50;  /* largest -dz/dr along the direction */
          for (double r = 1.0; r < q; r += 1.0) {
            const double rr = r * r * scale;
            const double xx = x + rr * cs[a], yy = y + rr * sn[a];
            if ((int)xx == x && (int)yy == y) continue;
            if (xx < 0 || xx > width - 1 || yy < 0 || yy > height - 1) continue;
            const double z2 = zb[(size_t)(int)yy * width + (int)xx];
            double xx2, yy2;
            ssao_sample(type, fov, aspect, width, height, xx, yy, z2, &xx2, &yy2);
            const double dx = xx2 - x2, dy = yy2 - y2, dz = z2 - z;
            const double dr = sqrt(dx * dx + dy * dy);
            const double diff = -dz / dr;
            if (diff > best) best = diff;
          }
          ambient += -atan(best) / M_PI + 0.5;
        }
        total = ambient / q;
        if (total < 0) total = 0;
      }
      for (int c = 0; c < 3; c++) {  /* image.cpp:375-383 */
        double v = image16[3 * p + c] + 65535.0 * colour[c][p] / 256.0 * total * intensity * (1.0 - op);
        if (v > 65535) v = 65535;
        image16[3 * p + c] = (uint16_t)v;
      }
    }
  }
  free(sn);
  free(cs);
  return MBR_OK;
}

/* ----------------------------------------------------------------- DOF */

typedef struct { float z; int i; } zsort_t;

/* QuickSortZBuffer (src/common_math.h:53-84): Hoare partition around the
 * middle element, ascending z; recursion on the two sides. */
static void zsort(zsort_t *d, int l, int p) {
  int i = l, j = p;
  const zsort_t x = d[(l + p) / 2];
  do {
    while (d[i].z < x.z) i++;
    while (x.z < d[j].z) j--;
    if (i <= j) {
      zsort_t t = d[i]; d[i] = d[j]; d[j] = t;
      i++; j--;
    }
  } while (i <= j);
  if (l < j) zsort(d, l, j);
  if (i < p) zsort(d, i, p);
}

int oracle_dof_order(const mbr_scene *s, int width, int height, const float *zb, uint32_t seed, int *order) {
  const int n = width * height;
  if (n <= 0) return MBR_ERR_INVALID_ARG;
  const double deep = s->dof_radius * width / 1000.0;               /* Render3D.cpp:1103-1104 */
  const double neutral = pow(10, s->dof_focus / 10.0 - 16.0);
  zsort_t *d = (zsort_t *)malloc(sizeof(zsort_t) * n);
  for (int k = 0; k < n; k++) { d[k].z = zb[k]; d[k].i = k; }
  if (n > 1) zsort(d, 1, n - 1);                                     /* image.cpp:86: from index 1 */
  srand(seed);
  for (int i = n - 1; i >= 0; i--) {                                 /* image.cpp:96-152 */
    const zsort_t t = d[i];
    const double z1 = t.z;
    const double size1 = (z1 - neutral) / z1 * deep;
    int step = i, ii, done;
    do {
      ii = i - (int)(rand() % (step + 1));
      if (ii <= 0) ii = 0;
      const double z2 = d[ii].z;
      const double size2 = (z2 - neutral) / z2 * deep;
      if (size1 * size2 > 0) {
        const double cmp = size1 > 0 ? size2 / size1 : size1 / size2;
        done = cmp > 0.7;
      } else {
        done = 0;  /* also when both sizes are 0: the assignment above it is overwritten */
      }
      step = (int)(step * 0.7 - 1);
      if (step <= 0) done = 1;
    } while (!done);
    d[i] = d[ii];
    d[ii] = t;
  }
  for (int k = 0; k < n; k++) order[k] = d[n - 1 - k].i;            /* processed from the end */
  free(d);
  return MBR_OK;
}

int oracle_dof(const mbr_scene *s, int width, int height, const float *zb, uint32_t seed, uint16_t *image16,
               uint16_t *alpha) {
  const int n = width * height;
  if (n <= 0) return MBR_ERR_INVALID_ARG;
  int *order = (int *)malloc(sizeof(int) * n);
  int rc = oracle_dof_order(s, width, height, zb, seed, order);
  if (rc != MBR_OK) { free(order); return rc; }
  const double deep = s->dof_radius * width / 1000.0;
  const double neutral = pow(10, s->dof_focus / 10.0 - 16.0);
  uint16_t *img0 = (uint16_t *)malloc(sizeof(uint16_t) * 3 * (size_t)n);
  uint16_t *a0 = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)n);
  memcpy(img0, image16, sizeof(uint16_t) * 3 * (size_t)n);
  if (alpha) memcpy(a0, alpha, sizeof(uint16_t) * (size_t)n);
  for (int k = 0; k < n; k++) {                                      /* image.cpp:165-205 */
    const int src = order[k];
    const int x = src % width, y = src / width;
    const double z = zb[src];
    double blur = fabs(z - neutral) / z * deep;
    if (blur > 100) blur = 100.0;
    const int size = (int)blur;
    const double factor = blur * blur * sqrt(blur) * M_PI / 3.0;
    for (int yy = y - size; yy <= y + size; yy++) {
      for (int xx = x - size; xx <= x + size; xx++) {
        if (xx < 0 || xx >= width || yy < 0 || yy >= height) continue;
        const int dx = xx - x, dy = yy - y;
        const double r = sqrtf((float)dx * dx + dy * dy);           /* the float overload in C++ */
        double op = (blur - r) / factor;
        if (op > 1.0) op = 1.0;
        if (op < 0.0) op = 0.0;
        if (op > 0.0) {
          const double opN = 1.0 - op;
          const size_t q = (size_t)yy * width + xx;
          for (int c = 0; c < 3; c++) image16[3 * q + c] = (uint16_t)(image16[3 * q + c] * opN + img0[3 * (size_t)src + c] * op);
          if (alpha) alpha[q] = (uint16_t)(alpha[q] * opN + a0[src] * op);
        }
      }
    }
  }
  free(order); free(img0); free(a0);
  return MBR_OK;
}
