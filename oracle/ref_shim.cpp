// ref_shim.cpp — C entry points over the reference's own algebra.cpp,
// common_math.cpp and primitives.cpp (compiled from /root/reference/src by
// oracle/Makefile into oracle/_ref/libmbref.so).  TEST INFRASTRUCTURE ONLY:
// it pins the oracle's camera, projection, primitive and Random()
// conventions against the compiled reference arithmetic.
#include <cstdlib>
#include "algebra.hpp"
#include "common_math.h"
#include "primitives.h"

static void get_m(const CRotationMatrix &r, double *m) {
  // images of the base vectors, i.e. the columns of the matrix
  CVector3 ex = r.RotateVector(CVector3(1, 0, 0));
  CVector3 ey = r.RotateVector(CVector3(0, 1, 0));
  CVector3 ez = r.RotateVector(CVector3(0, 0, 1));
  m[0] = ex.x; m[1] = ey.x; m[2] = ez.x;
  m[3] = ex.y; m[4] = ey.y; m[5] = ez.y;
  m[6] = ex.z; m[7] = ey.z; m[8] = ez.z;
}

extern "C" {

// camera matrix as MainThread builds it (src/Render3D.cpp:237-240)
void ref_camera_matrix(double alpha, double beta, double gamma, double *m) {
  CRotationMatrix r;
  r.RotateZ(alpha);
  r.RotateX(beta);
  r.RotateY(gamma);
  get_m(r, m);
}

// RotateVector of the camera matrix
void ref_camera_rotate(double alpha, double beta, double gamma, const double *v, double *out) {
  CRotationMatrix r;
  r.RotateZ(alpha);
  r.RotateX(beta);
  r.RotateY(gamma);
  CVector3 o = r.RotateVector(CVector3(v[0], v[1], v[2]));
  out[0] = o.x; out[1] = o.y; out[2] = o.z;
}

// CRotationMatrix::SetRotation(angles[3]) (Mandelbox rotations)
void ref_set_rotation3(const double *angles, double *m) {
  CRotationMatrix r;
  double a[3] = {angles[0], angles[1], angles[2]};
  r.SetRotation(a);
  get_m(r, m);
}

// InvProjection3D (src/common_math.cpp:111-150): world point -> screen
void ref_inv_projection(const double *point, const double *vp, double alpha, double beta, double gamma,
                        int perspective_type, double fov, double zoom, double w, double h, double *out) {
  CRotationMatrix r;
  r.RotateZ(alpha);
  r.RotateX(beta);
  r.RotateY(gamma);
  CRotationMatrix rinv = r.Transpose();
  CVector3 s = InvProjection3D(CVector3(point[0], point[1], point[2]), CVector3(vp[0], vp[1], vp[2]), rinv,
                               (enumPerspectiveType)perspective_type, fov, zoom, w, h);
  out[0] = s.x; out[1] = s.y; out[2] = s.z;
}

double ref_primitive_plane(const double *p, const double *c, const double *n) {
  return PrimitivePlane(CVector3(p[0], p[1], p[2]), CVector3(c[0], c[1], c[2]), CVector3(n[0], n[1], n[2]));
}
double ref_primitive_box(const double *p, const double *c, const double *s, int inverted) {
  CVector3 P(p[0], p[1], p[2]), C(c[0], c[1], c[2]), S(s[0], s[1], s[2]);
  return inverted ? PrimitiveInvertedBox(P, C, S) : PrimitiveBox(P, C, S);
}
double ref_primitive_sphere(const double *p, const double *c, double radius, int inverted) {
  CVector3 P(p[0], p[1], p[2]), C(c[0], c[1], c[2]);
  return inverted ? PrimitiveInvertedSphere(P, C, radius) : PrimitiveSphere(P, C, radius);
}
double ref_primitive_water(const double *p, double height, double amplitude, double length, double rotation,
                           int iterations, double anim_speed, int frame) {
  return PrimitiveWater(CVector3(p[0], p[1], p[2]), height, amplitude, length, rotation, iterations, anim_speed,
                        frame);
}

// srand(seed) then n x Random(max) (src/common_math.cpp:25-32)
void ref_random_sequence(unsigned seed, int max, int n, int *out) {
  srand(seed);
  for (int i = 0; i < n; i++) out[i] = Random(max);
}

double ref_dmax(double a, double b, double c) { return dMax(a, b, c); }
}
