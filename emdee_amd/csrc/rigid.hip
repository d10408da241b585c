// rigid.hip -- rigid bodies on the GPU (reference: src/ArBee.f90 and the
// body parts of src/EmDeeData.f90 move/boost/kinetic/virial/update).
//
// One thread per body.  Body state is SoA over bodies ([k][nb] for k-vector
// fields); members are stored in CSR order with their original atom index,
// and the current slot of a member is iperm[orig].  The per-body arithmetic
// follows the cited reference routines in FP64: Kopp's analytic 3x3
// eigensolver for the principal axes, Shepperd's quaternion, the exact
// torque-free rotation of an asymmetric top (Jacobi elliptic functions and
// Carlson integrals) and the NO_SQUISH splitting.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "device_state.h"
#include "engine.h"

namespace emdee {

constexpr double EPSD = 2.220446049250313e-16;   // epsilon(1.0_rb)
constexpr double PI_D = 3.14159265358979324;      // global.f90:38

// ---------------------------------------------------------------- math.f90

__device__ __forceinline__ double sgn(double a, double b) { return b >= 0 ? fabs(a) : -fabs(a); }
__device__ __forceinline__ void cross(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
// ArBee.f90:138-176; row-major 4x3 (B[3*r + c])
__device__ __forceinline__ void matB(const double* q, double* B) {
  B[0] = -q[1]; B[1] = -q[2]; B[2] = -q[3];
  B[3] = q[0];  B[4] = -q[3]; B[5] = q[2];
  B[6] = q[3];  B[7] = q[0];  B[8] = -q[1];
  B[9] = -q[2]; B[10] = q[1]; B[11] = q[0];
}
__device__ __forceinline__ void matC(const double* q, double* C) {
  C[0] = -q[1]; C[1] = -q[2]; C[2] = -q[3];
  C[3] = q[0];  C[4] = q[3];  C[5] = -q[2];
  C[6] = -q[3]; C[7] = q[0];  C[8] = q[1];
  C[9] = q[2];  C[10] = -q[1]; C[11] = q[0];
}
__device__ __forceinline__ void mul43(const double* M, const double* x, double* y) {
  for (int r = 0; r < 4; r++) y[r] = M[3 * r] * x[0] + M[3 * r + 1] * x[1] + M[3 * r + 2] * x[2];
}
__device__ __forceinline__ void mul34t(const double* M, const double* x, double* y) {
  for (int c = 0; c < 3; c++) y[c] = M[c] * x[0] + M[3 + c] * x[1] + M[6 + c] * x[2] + M[9 + c] * x[3];
}

// Carlson's duplication algorithm for RF, RC, RJ: math.f90:505-638
__device__ double carlson_rf(double x, double y, double z) {
  const double errtol = 0.001, c1 = 1.0 / 24.0, c2 = 3.0 / 44.0, c3 = 1.0 / 14.0;
  double xn = x, yn = y, zn = z, mu, xd, yd, zd;
  for (int it = 0; it < 200; it++) {
    mu = (xn + yn + zn) / 3.0;
    xd = 2.0 - (mu + xn) / mu; yd = 2.0 - (mu + yn) / mu; zd = 2.0 - (mu + zn) / mu;
    if (fmax(fabs(xd), fmax(fabs(yd), fabs(zd))) < errtol) break;
    double xr = sqrt(xn), yr = sqrt(yn), zr = sqrt(zn);
    double lam = xr * (yr + zr) + yr * zr;
    xn = (xn + lam) * 0.25; yn = (yn + lam) * 0.25; zn = (zn + lam) * 0.25;
  }
  double e2 = xd * yd - zd * zd, e3 = xd * yd * zd;
  double s = 1.0 + (c1 * e2 - 0.10 - c2 * e3) * e2 + c3 * e3;
  return s / sqrt(mu);
}
__device__ double carlson_rc(double x, double y) {
  const double errtol = 0.001, c1 = 1.0 / 7.0, c2 = 9.0 / 22.0;
  double xn = x, yn = y, mu, sn;
  for (int it = 0; it < 200; it++) {
    mu = (xn + yn + yn) / 3.0;
    sn = (yn + mu) / mu - 2.0;
    if (fabs(sn) < errtol) break;
    double lam = 2.0 * sqrt(xn) * sqrt(yn) + yn;
    xn = (xn + lam) * 0.25; yn = (yn + lam) * 0.25;
  }
  double s = sn * sn * (0.30 + sn * (c1 + sn * (0.3750 + sn * c2)));
  return (1.0 + s) / sqrt(mu);
}
__device__ double carlson_rj(double x, double y, double z, double p) {
  const double errtol = 0.001, c1 = 3.0 / 14.0, c2 = 1.0 / 3.0, c3 = 3.0 / 22.0, c4 = 3.0 / 26.0;
  double xn = x, yn = y, zn = z, pn = p, sigma = 0.0, power4 = 1.0, mu, xd, yd, zd, pd;
  for (int it = 0; it < 200; it++) {
    mu = (xn + yn + zn + pn + pn) * 0.20;
    xd = (mu - xn) / mu; yd = (mu - yn) / mu; zd = (mu - zn) / mu; pd = (mu - pn) / mu;
    if (fmax(fmax(fabs(xd), fabs(yd)), fmax(fabs(zd), fabs(pd))) < errtol) break;
    double xr = sqrt(xn), yr = sqrt(yn), zr = sqrt(zn);
    double lam = xr * (yr + zr) + yr * zr;
    double alfa = pn * (xr + yr + zr) + xr * yr * zr;
    alfa = alfa * alfa;
    double beta = pn * (pn + lam) * (pn + lam);
    sigma = sigma + power4 * carlson_rc(alfa, beta);
    power4 = power4 * 0.25;
    xn = (xn + lam) * 0.25; yn = (yn + lam) * 0.25; zn = (zn + lam) * 0.25; pn = (pn + lam) * 0.25;
  }
  double ea = xd * (yd + zd) + yd * zd, eb = xd * yd * zd, ec = pd * pd;
  double e2 = ea - 3.0 * ec, e3 = eb + 2.0 * pd * (ea - ec);
  double s1 = 1.0 + e2 * (-c1 + 0.750 * c3 * e2 - 1.50 * c4 * e3);
  double s2 = eb * (0.50 * c2 + pd * (-c3 - c3 + pd * c4));
  double s3 = pd * ea * (c2 - pd * c3) - c2 * pd * ec;
  return 3.0 * sigma + power4 * (s1 + s2 + s3) / (mu * sqrt(mu));
}

// Jacobi elliptic sn, cn, dn by the arithmetic-geometric mean: math.f90:440-501
__device__ void jacobi(double u, double m, double* jac) {
  const int NN = 16;
  double mu[NN], nu[NN], c[NN], d[NN];
  if (fabs(m) > 1.0) { jac[0] = jac[1] = jac[2] = NAN; return; }
  if (fabs(m) < 2.0 * EPSD) { jac[0] = sin(u); jac[1] = cos(u); jac[2] = 1.0; return; }
  if (fabs(m - 1.0) < 2.0 * EPSD) { jac[0] = tanh(u); jac[1] = jac[2] = 1.0 / cosh(u); return; }
  int n = 0;
  mu[0] = 1.0;
  nu[0] = sqrt(1.0 - m);
  while (fabs(mu[n] - nu[n]) > 4.0 * EPSD * fabs(mu[n] + nu[n])) {
    mu[n + 1] = 0.5 * (mu[n] + nu[n]);
    nu[n + 1] = sqrt(mu[n] * nu[n]);
    n = n + 1;
    if (n >= NN - 1) { jac[0] = jac[1] = jac[2] = NAN; return; }
  }
  double s = sin(u * mu[n]), co = cos(u * mu[n]), r;
  if (fabs(s) < fabs(co)) {
    c[n] = mu[n] * (s / co);
    d[n] = 1.0;
    while (n > 0) {
      c[n - 1] = d[n] * c[n];
      r = (c[n] * c[n]) / mu[n];
      n = n - 1;
      d[n] = (r + nu[n]) / (r + mu[n]);
    }
    double dn = sqrt(1.0 - m) / d[n];
    double cn = dn * sgn(1.0, co) / hypot(1.0, c[n]);
    jac[0] = cn * c[n] / sqrt(1.0 - m); jac[1] = cn; jac[2] = dn;
  } else {
    c[n] = mu[n] * (co / s);
    d[n] = 1.0;
    while (n > 0) {
      c[n - 1] = d[n] * c[n];
      r = (c[n] * c[n]) / mu[n];
      n = n - 1;
      d[n] = (r + nu[n]) / (r + mu[n]);
    }
    double sn = sgn(1.0, s) / hypot(1.0, c[n]);
    jac[0] = sn; jac[1] = c[n] * sn; jac[2] = d[n];
  }
}

__device__ __forceinline__ int staircase(double x) {   // math.f90:428-436
  return x > 0.0 ? (int)ceil(x - 0.5) : (int)floor(x + 0.5);
}

// Kopp's analytic eigensolver (Cardano + cross products): math.f90:244-424.
// a is column-major a[(r-1)+3(c-1)]; returns eigenvectors as columns of q.
#define AE(A, r, c) A[((r) - 1) + 3 * ((c) - 1)]
__device__ void eigvec(double* v, const double* a, double n1t, double n2t, double thresh) {
  double norm = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  double n1 = n1t + AE(a, 1, 1) * AE(a, 1, 1), n2 = n2t + AE(a, 2, 2) * AE(a, 2, 2), err = n1 * n2;
  if (n1 <= thresh) { v[0] = 1.0; v[1] = 0.0; v[2] = 0.0; }
  else if (n2 <= thresh) { v[0] = 0.0; v[1] = 1.0; v[2] = 0.0; }
  else if (norm < (64.0 * EPSD) * (64.0 * EPSD) * err) {
    double t = fabs(AE(a, 1, 2)), f = -AE(a, 1, 1) / AE(a, 1, 2);
    if (fabs(AE(a, 2, 2)) > t) { t = fabs(AE(a, 2, 2)); f = -AE(a, 1, 2) / AE(a, 2, 2); }
    if (fabs(AE(a, 2, 3)) > t) f = -AE(a, 1, 3) / AE(a, 2, 3);
    norm = 1.0 / sqrt(1.0 + f * f);
    v[0] = norm; v[1] = f * norm; v[2] = 0.0;
  } else {
    double s = sqrt(1.0 / norm);
    v[0] *= s; v[1] *= s; v[2] *= s;
  }
}
__device__ void diagonalize(const double* m, double* q, double* w) {
  double a[9];
  for (int k = 0; k < 9; k++) a[k] = m[k];
  double de = AE(a, 1, 2) * AE(a, 2, 3), dd = AE(a, 1, 2) * AE(a, 1, 2), ee = AE(a, 2, 3) * AE(a, 2, 3);
  double ff = AE(a, 1, 3) * AE(a, 1, 3), mm = AE(a, 1, 1) + AE(a, 2, 2) + AE(a, 3, 3);
  double c1 = (AE(a, 1, 1) * AE(a, 2, 2) + AE(a, 1, 1) * AE(a, 3, 3) + AE(a, 2, 2) * AE(a, 3, 3)) - (dd + ee + ff);
  double c0 = 27.0 * (AE(a, 3, 3) * dd + AE(a, 1, 1) * ee + AE(a, 2, 2) * ff - AE(a, 1, 1) * AE(a, 2, 2) * AE(a, 3, 3) -
                      2.0 * AE(a, 1, 3) * de);
  double p = mm * mm - 3.0 * c1, r = mm * (p - 1.5 * c1) - 0.5 * c0, sqrtp = sqrt(fabs(p));
  double phi = (1.0 / 3.0) * atan2(sqrt(fabs(6.75 * c1 * c1 * (p - c1) + c0 * (r + 0.25 * c0))), r);
  double c = sqrtp * cos(phi), s = (1.0 / sqrt(3.0)) * sqrtp * sin(phi);
  p = (1.0 / 3.0) * (mm - c);
  double w1 = p + c, w2 = p - s, w3 = p + s, t;
  if (fabs(w1) < fabs(w3)) { t = w1; w1 = w3; w3 = t; }
  if (fabs(w1) < fabs(w2)) { t = w1; w1 = w2; w2 = t; }
  if (fabs(w2) < fabs(w3)) { t = w2; w2 = w3; w3 = t; }
  w[0] = w1; w[1] = w2; w[2] = w3;
  double wmax8eps = 8.0 * EPSD * fabs(w1), thresh = wmax8eps * wmax8eps;
  double n1 = AE(a, 1, 2) * AE(a, 1, 2) + AE(a, 1, 3) * AE(a, 1, 3);
  double n2 = AE(a, 1, 2) * AE(a, 1, 2) + AE(a, 2, 3) * AE(a, 2, 3);
  double q11 = AE(a, 1, 2) * AE(a, 2, 3) - AE(a, 1, 3) * AE(a, 2, 2);
  double q21 = AE(a, 1, 3) * AE(a, 1, 2) - AE(a, 2, 3) * AE(a, 1, 1);
  double q32 = AE(a, 1, 2) * AE(a, 1, 2);
  AE(a, 1, 1) -= w1; AE(a, 2, 2) -= w1;
  double v[3] = { q11 + AE(a, 1, 3) * w1, q21 + AE(a, 2, 3) * w1, AE(a, 1, 1) * AE(a, 2, 2) - q32 };
  eigvec(v, a, n1, n2, thresh);
  AE(q, 1, 1) = v[0]; AE(q, 2, 1) = v[1]; AE(q, 3, 1) = v[2];
  t = w1 - w2;
  if (fabs(t) > wmax8eps) {
    AE(a, 1, 1) += t; AE(a, 2, 2) += t;
    double u[3] = { q11 + AE(a, 1, 3) * w2, q21 + AE(a, 2, 3) * w2, AE(a, 1, 1) * AE(a, 2, 2) - q32 };
    eigvec(u, a, n1, n2, thresh);
    AE(q, 1, 2) = u[0]; AE(q, 2, 2) = u[1]; AE(q, 3, 2) = u[2];
  } else {
    AE(a, 2, 1) = AE(a, 1, 2); AE(a, 3, 1) = AE(a, 1, 3); AE(a, 3, 2) = AE(a, 2, 3);
    AE(a, 1, 1) += w1; AE(a, 2, 2) += w1;
    int i = 0;
    bool success = false;
    double u[3] = { 0, 0, 0 };
    while (i < 3 && !success) {
      i = i + 1;
      AE(a, i, i) -= w2;
      double nn = AE(a, 1, i) * AE(a, 1, i) + AE(a, 2, i) * AE(a, 2, i) + AE(a, 3, i) * AE(a, 3, i);
      success = nn > thresh;
      if (success) {
        double q1[3] = { AE(q, 1, 1), AE(q, 2, 1), AE(q, 3, 1) }, ai[3] = { AE(a, 1, i), AE(a, 2, i), AE(a, 3, i) };
        cross(q1, ai, u);
        double norm = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        success = norm > (256.0 * EPSD) * (256.0 * EPSD) * nn;
        if (success) { double f = sqrt(1.0 / norm); u[0] *= f; u[1] *= f; u[2] *= f; }
        AE(q, 1, 2) = u[0]; AE(q, 2, 2) = u[1]; AE(q, 3, 2) = u[2];
      }
    }
    if (!success) {
      int ii = 1;
      while (AE(q, ii, 1) == 0.0) ii = ii + 1;
      int j = 1 + ii % 3;
      double norm = 1.0 / sqrt(AE(q, ii, 1) * AE(q, ii, 1) + AE(q, j, 1) * AE(q, j, 1));
      AE(q, ii, 2) = AE(q, j, 1) * norm;
      AE(q, j, 2) = -AE(q, ii, 1) * norm;
      AE(q, 1 + (ii + 1) % 3, 2) = 0.0;
    }
  }
  double q1[3] = { AE(q, 1, 1), AE(q, 2, 1), AE(q, 3, 1) }, q2[3] = { AE(q, 1, 2), AE(q, 2, 2), AE(q, 3, 2) }, q3[3];
  cross(q1, q2, q3);
  AE(q, 1, 3) = q3[0]; AE(q, 2, 3) = q3[1]; AE(q, 3, 3) = q3[2];
}
// Shepperd's quaternion of a rotation matrix: math.f90:189-218
__device__ void quaternion_of(const double* A, double* Q) {
  double a11 = AE(A, 1, 1), a22 = AE(A, 2, 2), a33 = AE(A, 3, 3);
  double Q2[4] = { 1.0 + a11 + a22 + a33, 1.0 + a11 - a22 - a33, 1.0 - a11 + a22 - a33, 1.0 - a11 - a22 + a33 };
  int im = 0;
  for (int k = 1; k < 4; k++) if (Q2[k] > Q2[im]) im = k;
  double Qm = Q2[im], f = sqrt(1.0 / Qm), v[4];
  if (im == 0) { v[0] = Qm; v[1] = AE(A,2,3) - AE(A,3,2); v[2] = AE(A,3,1) - AE(A,1,3); v[3] = AE(A,1,2) - AE(A,2,1); }
  else if (im == 1) { v[0] = AE(A,2,3) - AE(A,3,2); v[1] = Qm; v[2] = AE(A,1,2) + AE(A,2,1); v[3] = AE(A,1,3) + AE(A,3,1); }
  else if (im == 2) { v[0] = AE(A,3,1) - AE(A,1,3); v[1] = AE(A,1,2) + AE(A,2,1); v[2] = Qm; v[3] = AE(A,2,3) + AE(A,3,2); }
  else { v[0] = AE(A,1,2) - AE(A,2,1); v[1] = AE(A,1,3) + AE(A,3,1); v[2] = AE(A,2,3) + AE(A,3,2); v[3] = Qm; }
  for (int k = 0; k < 4; k++) Q[k] = 0.5 * v[k] * f;
}

// ----------------------------------------------------------------- ArBee.f90

struct BodyReg {
  double MoI[3], invMoI[3], I113, I313, I223, I212, m312;
  double q[4], pi[4], omega[3];
};

__device__ __forceinline__ void load_body(const BodyState& B, int b, BodyReg& r) {
  int nb = B.nb;
  for (int k = 0; k < 3; k++) { r.MoI[k] = B.MoI.p[k * nb + b]; r.invMoI[k] = B.invMoI.p[k * nb + b]; r.omega[k] = B.omega.p[k * nb + b]; }
  for (int k = 0; k < 4; k++) { r.q[k] = B.q.p[k * nb + b]; r.pi[k] = B.pi.p[k * nb + b]; }
  r.I113 = B.coef.p[b]; r.I313 = B.coef.p[nb + b]; r.I223 = B.coef.p[2 * nb + b]; r.I212 = B.coef.p[3 * nb + b];
  r.m312 = B.coef.p[4 * nb + b];
}
__device__ __forceinline__ void store_motion(const BodyState& B, int b, const BodyReg& r) {
  int nb = B.nb;
  for (int k = 0; k < 3; k++) B.omega.p[k * nb + b] = r.omega[k];
  for (int k = 0; k < 4; k++) { B.q.p[k * nb + b] = r.q[k]; B.pi.p[k * nb + b] = r.pi[k]; }
}

// tBody_assign_momenta, ArBee.f90:356-366 (quaternion momenta given)
__device__ __forceinline__ void assign_pi(BodyReg& r, const double* pin) {
  double B[12], t[3];
  for (int k = 0; k < 4; k++) r.pi[k] = pin[k];
  matB(r.q, B);
  mul34t(B, r.pi, t);
  for (int k = 0; k < 3; k++) r.omega[k] = 0.5 * r.invMoI[k] * t[k];
}

// ArBee.f90:200-221
__device__ void rotate_uniaxial(BodyReg& b, int k, double dt) {
  double BkQ[4], BkPi[4];
  const double *q = b.q, *p = b.pi;
  if (k == 1) {
    BkQ[0] = -q[1]; BkQ[1] = q[0]; BkQ[2] = q[3]; BkQ[3] = -q[2];
    BkPi[0] = -p[1]; BkPi[1] = p[0]; BkPi[2] = p[3]; BkPi[3] = -p[2];
  } else if (k == 2) {
    BkQ[0] = -q[2]; BkQ[1] = -q[3]; BkQ[2] = q[0]; BkQ[3] = q[1];
    BkPi[0] = -p[2]; BkPi[1] = -p[3]; BkPi[2] = p[0]; BkPi[3] = p[1];
  } else {
    BkQ[0] = -q[3]; BkQ[1] = q[2]; BkQ[2] = -q[1]; BkQ[3] = q[0];
    BkPi[0] = -p[3]; BkPi[1] = p[2]; BkPi[2] = -p[1]; BkPi[3] = p[0];
  }
  double om = dt * (p[0] * BkQ[0] + p[1] * BkQ[1] + p[2] * BkQ[2] + p[3] * BkQ[3]) / (4.0 * b.MoI[k - 1]);
  double vs = sin(om), vc = cos(om);
  for (int i = 0; i < 4; i++) { b.q[i] = vc * b.q[i] + vs * BkQ[i]; b.pi[i] = vc * b.pi[i] + vs * BkPi[i]; }
}
// ArBee.f90:180-196 (the final delta update is done by the caller)
__device__ void rotate_no_squish(BodyReg& b, double delta_t, int n) {
  double dt = delta_t / n, half_dt = 0.5 * dt;
  for (int i = 1; i <= n; i++) {
    rotate_uniaxial(b, 3, half_dt);
    rotate_uniaxial(b, 2, half_dt);
    rotate_uniaxial(b, 1, dt);
    rotate_uniaxial(b, 2, half_dt);
    rotate_uniaxial(b, 3, half_dt);
  }
}
__device__ __forceinline__ double theta_fn(double x, double n, double m) {   // ArBee.f90:292-297
  double x2 = x * x;
  return -(1.0 / 3.0) * n * x * x2 * carlson_rj(1.0 - x2, 1.0 - m * x2, 1.0, 1.0 + n * x2);
}
// Exact free rotation of an asymmetric top: ArBee.f90:225-324
__device__ void rotate_exact(BodyReg& b, double dt) {
  double w0[3], Iw[3];
  for (int k = 0; k < 3; k++) { w0[k] = b.omega[k]; Iw[k] = b.MoI[k] * w0[k]; }
  double Lsq = Iw[1] * Iw[1] + Iw[2] * Iw[2];
  if (Lsq < EPSD) { rotate_uniaxial(b, 1, dt); return; }
  Lsq = Iw[0] * Iw[0] + Lsq;
  double L = sqrt(Lsq), TwoKr = Iw[0] * w0[0] + Iw[1] * w0[1] + Iw[2] * w0[2];
  double r1 = Lsq - TwoKr * b.MoI[2], r3 = TwoKr * b.MoI[0] - Lsq;
  double l1 = b.I223 * r1, l3 = b.I212 * r3, lmin = fmin(l1, l3);
  double a[3] = { sgn(1.0, w0[0]) * sqrt(b.I113 * r1), sqrt(lmin), sgn(1.0, w0[2]) * sqrt(b.I313 * r3) };
  double m = lmin / fmax(l1, l3);
  double K = carlson_rf(0.0, 1.0 - m, 1.0), inv2K = 0.5 / K;
  double s0 = w0[1] / a[1], c0, u0;
  int i0;
  if (fabs(s0) < 1.0) {
    c0 = (l1 < l3) ? w0[0] / a[0] : w0[2] / a[2];
    u0 = s0 * carlson_rf(1.0 - s0 * s0, 1.0 - m * s0 * s0, 1.0);
    i0 = staircase(u0 * inv2K);
  } else {
    a[1] = fabs(w0[1]);
    s0 = sgn(1.0, s0);
    c0 = 0.0;
    u0 = sgn(K, s0);
    i0 = 0;
  }
  double wp = b.m312 * a[0] * a[2] / a[1];
  double u = wp * dt + u0;
  int jump = staircase(u * inv2K) - i0;
  double jac[3];
  jacobi(u, m, jac);
  double sn = jac[0], cn = jac[1], dn = jac[2], deltaF;
  double alpha = b.MoI[0] * a[0] / L;
  if (l1 < l3) {
    b.omega[0] = a[0] * cn; b.omega[1] = a[1] * sn; b.omega[2] = a[2] * dn;
    double d0 = w0[2] / a[2];
    double eta = alpha * alpha / (1.0 - alpha * alpha), C = sqrt(m + eta);
    double r = u - u0 + sgn(1.0, cn) * theta_fn(sn, eta, m) - sgn(1.0, c0) * theta_fn(s0, eta, m) +
               (alpha / C) * (atan(C * sn / dn) - atan(C * s0 / d0));
    if (jump != 0) r = r + jump * 2.0 * theta_fn(1.0, eta, m);
    deltaF = (eta + 1.0) * r;
  } else {
    b.omega[0] = a[0] * dn; b.omega[1] = a[1] * sn; b.omega[2] = a[2] * cn;
    double eta = alpha * alpha;
    eta = eta / (1.0 - eta);
    double k2eta = m * eta, C = sqrt(1.0 + k2eta);
    double r = u - u0 + sgn(1.0, cn) * theta_fn(sn, k2eta, m) - sgn(1.0, c0) * theta_fn(s0, k2eta, m) +
               (alpha / C) * (atan(C * sn / cn) - atan(C * s0 / c0));
    if (jump != 0) r = r + jump * (2.0 * theta_fn(1.0, k2eta, m) + (alpha / C) * PI_D);
    deltaF = (eta + 1.0) * r;
  }
  double phi = (Lsq * (u - u0) + r3 * deltaF) / (2.0 * L * b.MoI[0] * wp);
  double z0[4] = { Iw[2], Iw[1], L - Iw[0], 0.0 };
  for (int k = 0; k < 3; k++) Iw[k] = b.MoI[k] * b.omega[k];
  double cph = cos(phi), sph = sin(phi);
  double z[4] = { Iw[2] * cph - Iw[1] * sph, Iw[1] * cph + Iw[2] * sph, (L - Iw[0]) * cph, (L - Iw[0]) * sph };
  double Cz[12], Cz0[12], t3[3], t4[4], qn[4];
  matC(z, Cz); matC(z0, Cz0);
  mul34t(Cz0, b.q, t3);
  mul43(Cz, t3, t4);
  double dq = z0[0] * b.q[0] + z0[1] * b.q[1] + z0[2] * b.q[2] + z0[3] * b.q[3];
  for (int k = 0; k < 4; k++) qn[k] = z[k] * dq + t4[k];
  double nrm = sqrt(qn[0] * qn[0] + qn[1] * qn[1] + qn[2] * qn[2] + qn[3] * qn[3]);
  for (int k = 0; k < 4; k++) b.q[k] = qn[k] / nrm;
  double B[12], twoIw[3] = { 2.0 * Iw[0], 2.0 * Iw[1], 2.0 * Iw[2] };
  matB(b.q, B);
  mul43(B, twoIw, b.pi);
}

// rotate_exact over the two lanes h = 0, 1 of a body (lanes 2b, 2b+1 of one
// wavefront).  Its longest chain runs u0 -> Jacobi functions at u -> theta at
// sn; lane 0 takes that chain and the atan at sn, lane 1 takes K and theta
// and the atan at s0.  The lanes swap their values and both finish with the
// expressions of rotate_exact, in its order (theta at 1 only when the angle
// wraps, as there), so both hold the same new state.
__device__ __forceinline__ double swap_lane(double v) { return __shfl_xor(v, 1, 64); }
__device__ void rotate_exact2(BodyReg& b, double dt, int h) {
  double w0[3], Iw[3];
  for (int k = 0; k < 3; k++) { w0[k] = b.omega[k]; Iw[k] = b.MoI[k] * w0[k]; }
  double Lsq = Iw[1] * Iw[1] + Iw[2] * Iw[2];
  if (Lsq < EPSD) { rotate_uniaxial(b, 1, dt); return; }
  Lsq = Iw[0] * Iw[0] + Lsq;
  double L = sqrt(Lsq), TwoKr = Iw[0] * w0[0] + Iw[1] * w0[1] + Iw[2] * w0[2];
  double r1 = Lsq - TwoKr * b.MoI[2], r3 = TwoKr * b.MoI[0] - Lsq;
  double l1 = b.I223 * r1, l3 = b.I212 * r3, lmin = fmin(l1, l3);
  double a[3] = { sgn(1.0, w0[0]) * sqrt(b.I113 * r1), sqrt(lmin), sgn(1.0, w0[2]) * sqrt(b.I313 * r3) };
  double m = lmin / fmax(l1, l3);
  double s0 = w0[1] / a[1], c0, u0 = 0.0, K = 0.0;
  const bool inner = fabs(s0) < 1.0;
  if (h == 1 || !inner) K = carlson_rf(0.0, 1.0 - m, 1.0);
  if (inner) {
    c0 = (l1 < l3) ? w0[0] / a[0] : w0[2] / a[2];
    if (h == 0) u0 = s0 * carlson_rf(1.0 - s0 * s0, 1.0 - m * s0 * s0, 1.0);
  } else {
    a[1] = fabs(w0[1]);
    s0 = sgn(1.0, s0);
    c0 = 0.0;
    u0 = sgn(K, s0);
  }
  double wp = b.m312 * a[0] * a[2] / a[1];
  double alpha = b.MoI[0] * a[0] / L;
  const bool first = l1 < l3;
  double eta, tp, C, d0 = 0.0;
  if (first) {
    d0 = w0[2] / a[2];
    eta = alpha * alpha / (1.0 - alpha * alpha);
    tp = eta;
    C = sqrt(m + eta);
  } else {
    eta = alpha * alpha;
    eta = eta / (1.0 - eta);
    tp = m * eta;
    C = sqrt(1.0 + tp);
  }
  // lane 0: x = {u0, sn, cn, dn, theta(sn), atan at sn}; lane 1: x = {K,
  // theta(s0), atan at s0}
  double x[6] = { 0.0, 0.0, 0.0, 0.0, 0.0, 0.0 };
  if (h == 0) {
    double jac[3];
    jacobi(wp * dt + u0, m, jac);
    x[0] = u0; x[1] = jac[0]; x[2] = jac[1]; x[3] = jac[2];
    x[4] = theta_fn(jac[0], tp, m);
    x[5] = atan(C * jac[0] / (first ? jac[2] : jac[1]));
  } else {
    x[0] = K;
    x[1] = theta_fn(s0, tp, m);
    x[2] = atan(C * s0 / (first ? d0 : c0));
  }
  double y[6];
  for (int k = 0; k < 6; k++) y[k] = swap_lane(x[k]);
  const double* p0 = h == 0 ? x : y;   // lane 0's values
  const double* p1 = h == 0 ? y : x;   // lane 1's values
  if (inner) u0 = p0[0];
  K = p1[0];
  double inv2K = 0.5 / K;
  int i0 = inner ? staircase(u0 * inv2K) : 0;
  double u = wp * dt + u0;
  int jump = staircase(u * inv2K) - i0;
  double sn = p0[1], cn = p0[2], dn = p0[3], deltaF;
  if (first) {
    b.omega[0] = a[0] * cn; b.omega[1] = a[1] * sn; b.omega[2] = a[2] * dn;
    double r = u - u0 + sgn(1.0, cn) * p0[4] - sgn(1.0, c0) * p1[1] + (alpha / C) * (p0[5] - p1[2]);
    if (jump != 0) r = r + jump * 2.0 * theta_fn(1.0, tp, m);
    deltaF = (eta + 1.0) * r;
  } else {
    b.omega[0] = a[0] * dn; b.omega[1] = a[1] * sn; b.omega[2] = a[2] * cn;
    double r = u - u0 + sgn(1.0, cn) * p0[4] - sgn(1.0, c0) * p1[1] + (alpha / C) * (p0[5] - p1[2]);
    if (jump != 0) r = r + jump * (2.0 * theta_fn(1.0, tp, m) + (alpha / C) * PI_D);
    deltaF = (eta + 1.0) * r;
  }
  double phi = (Lsq * (u - u0) + r3 * deltaF) / (2.0 * L * b.MoI[0] * wp);
  double z0[4] = { Iw[2], Iw[1], L - Iw[0], 0.0 };
  for (int k = 0; k < 3; k++) Iw[k] = b.MoI[k] * b.omega[k];
  double cph = cos(phi), sph = sin(phi);
  double z[4] = { Iw[2] * cph - Iw[1] * sph, Iw[1] * cph + Iw[2] * sph, (L - Iw[0]) * cph, (L - Iw[0]) * sph };
  double Cz[12], Cz0[12], t3[3], t4[4], qn[4];
  matC(z, Cz); matC(z0, Cz0);
  mul34t(Cz0, b.q, t3);
  mul43(Cz, t3, t4);
  double dq = z0[0] * b.q[0] + z0[1] * b.q[1] + z0[2] * b.q[2] + z0[3] * b.q[3];
  for (int k = 0; k < 4; k++) qn[k] = z[k] * dq + t4[k];
  double nrm = sqrt(qn[0] * qn[0] + qn[1] * qn[1] + qn[2] * qn[2] + qn[3] * qn[3]);
  for (int k = 0; k < 4; k++) b.q[k] = qn[k] / nrm;
  double B[12], twoIw[3] = { 2.0 * Iw[0], 2.0 * Iw[1], 2.0 * Iw[2] };
  matB(b.q, B);
  mul43(B, twoIw, b.pi);
}

// delta_k = Ct(q) (B(q) d_k) for every member; R_k = rcm + delta_k when R given
__device__ void place_members(const BodyState& B, int b, const double* q, double* R, const int* iperm, int Npad) {
  double Bm[12], Cm[12];
  matB(q, Bm); matC(q, Cm);
  int nm = B.nm, nb = B.nb;
  double rc[3] = { B.rcm.p[b], B.rcm.p[nb + b], B.rcm.p[2 * nb + b] };
  for (int k = B.start.p[b]; k < B.start.p[b + 1]; k++) {
    double d[3] = { B.d.p[k], B.d.p[nm + k], B.d.p[2 * nm + k] }, t[4], dl[3];
    mul43(Bm, d, t);
    mul34t(Cm, t, dl);
    for (int x = 0; x < 3; x++) B.delta.p[x * nm + k] = dl[x];
    if (R) {
      int s = iperm[B.orig.p[k]];
      for (int x = 0; x < 3; x++) R[x * Npad + s] = rc[x] + dl[x];
    }
  }
}

// ------------------------------------------------------------------ kernels

// EmDeeData.f90:420-439 + ArBee.f90:97-134
__global__ void k_body_update(BodyState B, double* __restrict__ R, const int* __restrict__ iperm, double L, int Npad) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B.nb) return;
  int nb = B.nb, nm = B.nm, k0 = B.start.p[b], k1 = B.start.p[b + 1];
  double invL = 1.0 / L, r0[3];
  int s0 = iperm[B.orig.p[k0]];
  for (int x = 0; x < 3; x++) r0[x] = R[x * Npad + s0];
  // unwrap members next to the first one, in place
  for (int k = k0 + 1; k < k1; k++) {
    int s = iperm[B.orig.p[k]];
    for (int x = 0; x < 3; x++) {
      double v = R[x * Npad + s];
      R[x * Npad + s] = v - L * round(invL * (v - r0[x]));
    }
  }
  double invMass = B.invMass.p[b], rc[3];
  for (int x = 0; x < 3; x++) {
    double s = 0;
    for (int k = k0; k < k1; k++) s += B.mMass.p[k] * R[x * Npad + iperm[B.orig.p[k]]];
    rc[x] = s * invMass;
    B.rcm.p[x * nb + b] = rc[x];
  }
  double s11 = 0, s22 = 0, s33 = 0, s12 = 0, s13 = 0, s23 = 0;
  for (int k = k0; k < k1; k++) {
    int s = iperm[B.orig.p[k]];
    double d[3];
    for (int x = 0; x < 3; x++) { d[x] = R[x * Npad + s] - rc[x]; B.delta.p[x * nm + k] = d[x]; }
    double M = B.mMass.p[k];
    s11 += M * (d[1] * d[1] + d[2] * d[2]);
    s22 += M * (d[0] * d[0] + d[2] * d[2]);
    s33 += M * (d[0] * d[0] + d[1] * d[1]);
    s12 += M * d[0] * d[1];
    s13 += M * d[0] * d[2];
    s23 += M * d[1] * d[2];
  }
  double I[9] = { s11, 0, 0, -s12, s22, 0, -s13, -s23, s33 };   // column-major, upper triangle
  double A[9], At[9], w[3];
  diagonalize(I, A, w);
  for (int r = 1; r <= 3; r++) for (int c = 1; c <= 3; c++) AE(At, r, c) = AE(A, c, r);
  for (int x = 0; x < 3; x++) { B.MoI.p[x * nb + b] = w[x]; B.invMoI.p[x * nb + b] = 1.0 / w[x]; }
  double I1 = w[0], I2 = w[1], I3 = w[2];
  B.coef.p[b] = 1.0 / (I1 * (I1 - I3));
  B.coef.p[nb + b] = 1.0 / (I3 * (I1 - I3));
  B.coef.p[2 * nb + b] = 1.0 / (I2 * (I2 - I3));
  B.coef.p[3 * nb + b] = 1.0 / (I2 * (I1 - I2));
  B.coef.p[4 * nb + b] = (I3 - I1) / I2;
  double q[4];
  quaternion_of(At, q);
  for (int k = 0; k < 4; k++) B.q.p[k * nb + b] = q[k];
  for (int k = k0; k < k1; k++)
    for (int r = 1; r <= 3; r++)
      B.d.p[(r - 1) * nm + k] = AE(At, r, 1) * B.delta.p[k] + AE(At, r, 2) * B.delta.p[nm + k] + AE(At, r, 3) * B.delta.p[2 * nm + k];
}

// force_and_torque + boost for bodies (EmDeeData.f90:864-896, ArBee.f90:340-352),
// with the body parts of kinetic_energies (EmDeeData.f90:900-922).
template <bool KE>
__global__ void k_body_boost(BodyState B, const double* __restrict__ F, const int* __restrict__ iperm, double CP,
                             double CF, int translate, int rotate, int Npad) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  double v[6] = { 0, 0, 0, 0, 0, 0 };
  if (b < B.nb) {
    int nb = B.nb, nm = B.nm;
    double Fb[3] = { 0, 0, 0 }, tau[3] = { 0, 0, 0 };
    for (int k = B.start.p[b]; k < B.start.p[b + 1]; k++) {
      int s = iperm[B.orig.p[k]];
      double Fj[3] = { F[s], F[Npad + s], F[2 * Npad + s] };
      double dl[3] = { B.delta.p[k], B.delta.p[nm + k], B.delta.p[2 * nm + k] }, cr[3];
      for (int x = 0; x < 3; x++) Fb[x] = Fb[x] + Fj[x];
      cross(dl, Fj, cr);
      for (int x = 0; x < 3; x++) tau[x] = tau[x] + cr[x];
    }
    for (int x = 0; x < 3; x++) { B.F.p[x * nb + b] = Fb[x]; B.tau.p[x * nb + b] = tau[x]; }
    if (translate)
      for (int x = 0; x < 3; x++) B.pcm.p[x * nb + b] = CP * B.pcm.p[x * nb + b] + CF * Fb[x];
    BodyReg r;
    load_body(B, b, r);
    if (rotate) {
      double C[12], t[3] = { 2.0 * CF * tau[0], 2.0 * CF * tau[1], 2.0 * CF * tau[2] }, c4[4], in[4];
      matC(r.q, C);
      mul43(C, t, c4);
      for (int k = 0; k < 4; k++) in[k] = CP * r.pi[k] + c4[k];
      assign_pi(r, in);
      store_motion(B, b, r);
    }
    if (KE) {
      double im = B.invMass.p[b];
      for (int x = 0; x < 3; x++) {
        double p = B.pcm.p[x * nb + b];
        v[x] = im * (p * p);
        v[3 + x] = r.MoI[x] * (r.omega[x] * r.omega[x]);
      }
    }
  }
  if (KE) {
    block_sum<6>(v);
    if (threadIdx.x == 0) for (int k = 0; k < 6; k++) B.part.p[6 * blockIdx.x + k] = v[k];
  }
}

// move for bodies (EmDeeData.f90:823-860): translation of the centre of mass,
// then rotation and member positions.  Member positions are written only by
// the rotation branch, as in the reference.
// (256-thread launch bounds: the rotation's long FP64 chains keep their
// values in registers instead of spilling at the default 1024-thread bound)
// Two lanes per body (lanes 2b, 2b+1): the exact rotation splits its chain
// over them (rotate_exact2); lane 0 stores.
__global__ void __launch_bounds__(BLOCK) k_body_move(BodyState B, double* __restrict__ R, const int* __restrict__ iperm, double CR, double CP,
                            double dt, int translate, int rotate, int mode, int Npad) {
  const int gt = blockIdx.x * blockDim.x + threadIdx.x, b = gt >> 1, h = gt & 1;
  if (b >= B.nb) return;   // both lanes of a body
  int nb = B.nb;
  if (translate && h == 0)
    for (int x = 0; x < 3; x++)
      B.rcm.p[x * nb + b] = CR * B.rcm.p[x * nb + b] + CP * B.invMass.p[b] * B.pcm.p[x * nb + b];
  if (rotate) {
    BodyReg r;
    load_body(B, b, r);
    if (mode == 0) rotate_exact2(r, dt, h);
    else if (h == 0) rotate_no_squish(r, dt, mode);
    if (h == 0) {
      store_motion(B, b, r);
      place_members(B, b, r.q, R, iperm, Npad);
    }
  }
}

// rigid_body_virial partials: EmDeeData.f90:926-953
__global__ void k_body_virial(BodyState B, const double* __restrict__ F, const int* __restrict__ iperm, int Npad) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  double v[1] = { 0 };
  if (b < B.nb) {
    int nm = B.nm;
    for (int k = B.start.p[b]; k < B.start.p[b + 1]; k++) {
      int s = iperm[B.orig.p[k]];
      v[0] += F[s] * B.delta.p[k] + F[Npad + s] * B.delta.p[nm + k] + F[2 * Npad + s] * B.delta.p[2 * nm + k];
    }
  }
  block_sum<1>(v);
  if (threadIdx.x == 0) B.part.p[blockIdx.x] = v[0];
}

// assign_momenta for bodies from particle momenta (EmDeeData.f90:157-189);
// Paos: host momenta in original order, double[N][3].
__global__ void k_body_assign(BodyState B, const double* __restrict__ Paos) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  double v[6] = { 0, 0, 0, 0, 0, 0 };
  if (b < B.nb) {
    int nb = B.nb, nm = B.nm;
    double pcm[3] = { 0, 0, 0 }, Lv[3] = { 0, 0, 0 };
    for (int k = B.start.p[b]; k < B.start.p[b + 1]; k++) {
      const double* Pj = Paos + 3 * (size_t)B.orig.p[k];
      double dl[3] = { B.delta.p[k], B.delta.p[nm + k], B.delta.p[2 * nm + k] }, cr[3];
      for (int x = 0; x < 3; x++) pcm[x] = pcm[x] + Pj[x];
      cross(dl, Pj, cr);
      for (int x = 0; x < 3; x++) Lv[x] = Lv[x] + cr[x];
    }
    BodyReg r;
    load_body(B, b, r);
    double C[12], t[3] = { 2.0 * Lv[0], 2.0 * Lv[1], 2.0 * Lv[2] }, pin[4];
    matC(r.q, C);
    mul43(C, t, pin);
    assign_pi(r, pin);
    store_motion(B, b, r);
    double im = B.invMass.p[b];
    for (int x = 0; x < 3; x++) {
      B.pcm.p[x * nb + b] = pcm[x];
      v[x] = im * (pcm[x] * pcm[x]);
      v[3 + x] = r.MoI[x] * (r.omega[x] * r.omega[x]);
    }
  }
  block_sum<6>(v);
  if (threadIdx.x == 0) for (int k = 0; k < 6; k++) B.part.p[6 * blockIdx.x + k] = v[k];
}

// tBody_particle_momenta, ArBee.f90:328-336, into AoS original order
__global__ void k_body_particle_momenta(BodyState B, double* __restrict__ Paos) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B.nb) return;
  int nb = B.nb, nm = B.nm;
  double q[4], Bm[12], Cm[12], At[9], w[3], om[3];
  for (int k = 0; k < 4; k++) q[k] = B.q.p[k * nb + b];
  for (int k = 0; k < 3; k++) w[k] = B.omega.p[k * nb + b];
  matB(q, Bm); matC(q, Cm);
  for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) {
    double s = 0; for (int k = 0; k < 4; k++) s += Cm[3 * k + r] * Bm[3 * k + c];
    At[3 * r + c] = s;
  }
  for (int r = 0; r < 3; r++) om[r] = At[3 * r] * w[0] + At[3 * r + 1] * w[1] + At[3 * r + 2] * w[2];
  double im = B.invMass.p[b];
  for (int k = B.start.p[b]; k < B.start.p[b + 1]; k++) {
    double dl[3] = { B.delta.p[k], B.delta.p[nm + k], B.delta.p[2 * nm + k] }, cr[3];
    cross(om, dl, cr);
    double M = B.mMass.p[k];
    double* P = Paos + 3 * (size_t)B.orig.p[k];
    for (int x = 0; x < 3; x++) P[x] = M * (im * B.pcm.p[x * nb + b] + cr[x]);
  }
}

// member positions from rcm and q (after a host-side change of q or rcm)
__global__ void k_body_place(BodyState B, double* __restrict__ R, const int* __restrict__ iperm, int Npad) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B.nb) return;
  double q[4];
  for (int k = 0; k < 4; k++) q[k] = B.q.p[k * B.nb + b];
  place_members(B, b, q, R, iperm, Npad);
}

template <int NV>
__global__ void k_body_reduce(double* __restrict__ out, const double* __restrict__ part, int nblocks) {
  double v[NV];
  for (int k = 0; k < NV; k++) v[k] = 0;
#pragma unroll 8
  for (int b = threadIdx.x; b < nblocks; b += BLOCK)
    for (int k = 0; k < NV; k++) v[k] += part[NV * b + k];
  block_sum<NV>(v);
  if (threadIdx.x == 0) for (int k = 0; k < NV; k++) out[k] = v[k];
}

// ---- shadow terms of EmDee_verlet_step (EmDeeCode.f90:1107-1211) ----------

// virtual_rotation (EmDeeCode.f90:1150-1163): a copy of body b given the
// quaternion momenta pi + C(q) dt tau and rotated by dt; returns its q.
__device__ void virtual_rotation(const BodyState& B, int b, double dt, int mode, double* qout) {
  BodyReg r;
  load_body(B, b, r);
  int nb = B.nb;
  double C[12], t[3], c4[4], in[4];
  for (int x = 0; x < 3; x++) t[x] = dt * B.tau.p[x * nb + b];
  matC(r.q, C);
  mul43(C, t, c4);
  for (int k = 0; k < 4; k++) in[k] = r.pi[k] + c4[k];
  assign_pi(r, in);
  if (mode != 0) rotate_no_squish(r, dt, mode); else rotate_exact(r, dt);
  for (int k = 0; k < 4; k++) qout[k] = r.q[k];
}

// pre_force, body part: r0 = 2.5 rcm + dt/2 invMass (pcm - dt/2 F);
// q0 = q(-dt)/2 - 3 q.  sh holds [3][nb] r0 then [4][nb] q0.
__global__ void __launch_bounds__(BLOCK) k_body_shadow_pre(BodyState B, double* __restrict__ sh, double dt, int mode) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B.nb) return;
  int nb = B.nb;
  double dt_2 = 0.5 * dt, im = B.invMass.p[b], qv[4];
  for (int x = 0; x < 3; x++)
    sh[x * nb + b] = 2.5 * B.rcm.p[x * nb + b] + dt_2 * im * (B.pcm.p[x * nb + b] - dt_2 * B.F.p[x * nb + b]);
  virtual_rotation(B, b, -dt, mode, qv);
  for (int k = 0; k < 4; k++) sh[(3 + k) * nb + b] = 0.5 * qv[k] - 3.0 * B.q.p[k * nb + b];
}

// post_force, body part: partials of {Ks_t, Ks_r, Us}.
__global__ void __launch_bounds__(BLOCK) k_body_shadow_post(BodyState B, const double* __restrict__ sh, double dt, int mode) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  double v[3] = { 0.0, 0.0, 0.0 };
  if (b < B.nb) {
    int nb = B.nb;
    double dt_2 = 0.5 * dt, im = B.invMass.p[b];
    double q[4], pi[4], F[3], tau[3], pcm[3];
    for (int k = 0; k < 4; k++) { q[k] = B.q.p[k * nb + b]; pi[k] = B.pi.p[k * nb + b]; }
    for (int x = 0; x < 3; x++) { F[x] = B.F.p[x * nb + b]; tau[x] = B.tau.p[x * nb + b]; pcm[x] = B.pcm.p[x * nb + b]; }
    double rdot[3];
    for (int x = 0; x < 3; x++) rdot[x] = 2.5 * B.rcm.p[x * nb + b] + dt * im * (pcm[x] + dt_2 * F[x]) - sh[x * nb + b];
    v[0] = rdot[0] * pcm[0] + rdot[1] * pcm[1] + rdot[2] * pcm[2];
    double qv[4], qdot[4];
    virtual_rotation(B, b, dt, mode, qv);
    for (int k = 0; k < 4; k++) qdot[k] = sh[(3 + k) * nb + b] + 1.5 * q[k] + qv[k];
    double qq = qdot[0] * q[0] + qdot[1] * q[1] + qdot[2] * q[2] + qdot[3] * q[3], s = 0.0;
    for (int k = 0; k < 4; k++) s += (qdot[k] - qq * q[k]) * pi[k];
    v[1] = s;
    double Bm[12], C[12], c4[4], tb[3];
    matB(q, Bm);
    matC(q, C);
    mul43(C, tau, c4);
    mul34t(Bm, c4, tb);
    v[2] = im * (F[0] * F[0] + F[1] * F[1] + F[2] * F[2]) +
           (B.invMoI.p[b] * (tb[0] * tb[0]) + B.invMoI.p[nb + b] * (tb[1] * tb[1]) + B.invMoI.p[2 * nb + b] * (tb[2] * tb[2]));
  }
  block_sum<3>(v);
  if (threadIdx.x == 0) for (int k = 0; k < 3; k++) B.part.p[3 * blockIdx.x + k] = v[k];
}

// ------------------------------------------------------------ host side

static DBuf<double>* field(BodyState& B, int f, int& ncomp) {
  switch (f) {
    case BF_MASS: ncomp = 1; return &B.mass;
    case BF_MOI: ncomp = 3; return &B.MoI;
    case BF_INVMOI: ncomp = 3; return &B.invMoI;
    case BF_RCM: ncomp = 3; return &B.rcm;
    case BF_PCM: ncomp = 3; return &B.pcm;
    case BF_Q: ncomp = 4; return &B.q;
    case BF_PI: ncomp = 4; return &B.pi;
    case BF_OMEGA: ncomp = 3; return &B.omega;
    case BF_F: ncomp = 3; return &B.F;
    case BF_TAU: ncomp = 3; return &B.tau;
    default: fatal("rigid bodies", "invalid body field");
  }
}

void dev_bodies_setup(Device* d, const std::vector<int>& start, const std::vector<int>& orig,
                      const std::vector<double>& mMass, const std::vector<double>& bodyMass) {
  BodyState& B = d->bodies;
  B.nb = (int)bodyMass.size();
  B.nm = (int)orig.size();
  if (B.nb == 0) return;
  int nb = B.nb, nm = B.nm;
  B.start.alloc(nb + 1); B.orig.alloc(nm); B.mMass.alloc(nm); B.d.alloc(3 * (size_t)nm); B.delta.alloc(3 * (size_t)nm);
  B.mass.alloc(nb); B.invMass.alloc(nb); B.MoI.alloc(3 * (size_t)nb); B.invMoI.alloc(3 * (size_t)nb); B.coef.alloc(5 * (size_t)nb);
  B.rcm.alloc(3 * (size_t)nb); B.pcm.alloc(3 * (size_t)nb); B.q.alloc(4 * (size_t)nb); B.pi.alloc(4 * (size_t)nb);
  B.omega.alloc(3 * (size_t)nb); B.F.alloc(3 * (size_t)nb); B.tau.alloc(3 * (size_t)nb);
  B.part.alloc(6 * (size_t)(nb / BLOCK + 2));
  std::vector<double> im(nb);
  for (int b = 0; b < nb; b++) im[b] = 1.0 / bodyMass[b];
  HIP_CHECK(hipMemcpyAsync(B.start.p, start.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice, d->stream));
  HIP_CHECK(hipMemcpyAsync(B.orig.p, orig.data(), sizeof(int) * nm, hipMemcpyHostToDevice, d->stream));
  HIP_CHECK(hipMemcpyAsync(B.mMass.p, mMass.data(), sizeof(double) * nm, hipMemcpyHostToDevice, d->stream));
  HIP_CHECK(hipMemcpyAsync(B.mass.p, bodyMass.data(), sizeof(double) * nb, hipMemcpyHostToDevice, d->stream));
  HIP_CHECK(hipMemcpyAsync(B.invMass.p, im.data(), sizeof(double) * nb, hipMemcpyHostToDevice, d->stream));
  for (DBuf<double>* a : { &B.pcm, &B.pi, &B.omega, &B.F, &B.tau, &B.q, &B.rcm, &B.d, &B.delta, &B.MoI, &B.invMoI, &B.coef })
    HIP_CHECK(hipMemsetAsync(a->p, 0, sizeof(double) * a->n, d->stream));
  HIP_CHECK(hipStreamSynchronize(d->stream));
}

void dev_bodies_update(Device* d, double L) {
  d->posWrites++;
  BodyState& B = d->bodies;
  if (B.nb == 0) return;
  hipLaunchKernelGGL(k_body_update, dim3(div_up(B.nb, BLOCK)), dim3(BLOCK), 0, d->stream, B, d->R.p, d->iperm.p, L, d->Npad);
  launch_check("body update");
}

void dev_bodies_boost(Device* d, int layer, double CP, double CF, bool translate, bool rotate, double* twoKE6) {
  BodyState& B = d->bodies;
  if (B.nb == 0) { if (twoKE6) for (int k = 0; k < 6; k++) twoKE6[k] = 0; return; }
  int g = div_up(B.nb, BLOCK);
  const double* F = d->F.p + 3 * (size_t)d->Npad * layer;
  if (twoKE6) {
    hipLaunchKernelGGL(k_body_boost<true>, dim3(g), dim3(BLOCK), 0, d->stream, B, F, d->iperm.p, CP, CF, (int)translate, (int)rotate, d->Npad);
    // the totals straight into pinned host memory (same fixed-order sums)
    const unsigned long long seq = ++d->seq;
    hipLaunchKernelGGL(k_reduce_post<6>, dim3(1), dim3(BLOCK), 0, d->stream, d->hpostDev + POST_BODY_KE, d->hseqDev, seq,
                       B.part.p, g);
    dev_post_wait(d, seq);
    for (int k = 0; k < 6; k++) twoKE6[k] = d->hpost[POST_BODY_KE + k];
  } else {
    hipLaunchKernelGGL(k_body_boost<false>, dim3(g), dim3(BLOCK), 0, d->stream, B, F, d->iperm.p, CP, CF, (int)translate, (int)rotate, d->Npad);
  }
  launch_check("body boost");
}

void dev_bodies_move(Device* d, double CR, double CP, double dt, bool translate, bool rotate, int mode) {
  d->posWrites++;
  BodyState& B = d->bodies;
  if (B.nb == 0) return;
  hipLaunchKernelGGL(k_body_move, dim3(div_up(2L * B.nb, BLOCK)), dim3(BLOCK), 0, d->stream, B, d->R.p, d->iperm.p, CR, CP, dt,
                     (int)translate, (int)rotate, mode, d->Npad);
  launch_check("body move");
}

void dev_bodies_virial_launch(Device* d, int layer) {
  BodyState& B = d->bodies;
  if (B.nb == 0) { d->hpost[POST_BODY_W] = 0.0; return; }
  int g = div_up(B.nb, BLOCK);
  hipLaunchKernelGGL(k_body_virial, dim3(g), dim3(BLOCK), 0, d->stream, B, d->F.p + 3 * (size_t)d->Npad * layer, d->iperm.p, d->Npad);
  const unsigned long long seq = ++d->seq;
  hipLaunchKernelGGL(k_reduce_post<1>, dim3(1), dim3(BLOCK), 0, d->stream, d->hpostDev + POST_BODY_W, d->hseqDev, seq,
                     B.part.p, g);
}

double dev_bodies_virial_read(const Device* d) { return -d->hpost[POST_BODY_W]; }

double dev_bodies_virial(Device* d, int layer) {
  if (d->bodies.nb == 0) return 0.0;
  dev_bodies_virial_launch(d, layer);
  dev_wait_posts(d);
  return dev_bodies_virial_read(d);
}

void dev_bodies_assign_momenta(Device* d, double* twoKE6) {
  // the particle momenta are in d->stage (original order), uploaded by the caller
  BodyState& B = d->bodies;
  for (int k = 0; k < 6; k++) twoKE6[k] = 0;
  if (B.nb == 0) return;
  int g = div_up(B.nb, BLOCK);
  hipLaunchKernelGGL(k_body_assign, dim3(g), dim3(BLOCK), 0, d->stream, B, d->stage.p);
  const unsigned long long seq = ++d->seq;
  hipLaunchKernelGGL(k_reduce_post<6>, dim3(1), dim3(BLOCK), 0, d->stream, d->hpostDev + POST_BODY_KE, d->hseqDev, seq,
                     B.part.p, g);
  dev_post_wait(d, seq);
  for (int k = 0; k < 6; k++) twoKE6[k] = d->hpost[POST_BODY_KE + k];
}

void dev_bodies_particle_momenta(Device* d) {
  BodyState& B = d->bodies;
  if (B.nb == 0) return;
  hipLaunchKernelGGL(k_body_particle_momenta, dim3(div_up(B.nb, BLOCK)), dim3(BLOCK), 0, d->stream, B, d->stage.p);
  launch_check("body particle momenta");
}

void dev_bodies_place(Device* d) {
  d->posWrites++;
  BodyState& B = d->bodies;
  if (B.nb == 0) return;
  hipLaunchKernelGGL(k_body_place, dim3(div_up(B.nb, BLOCK)), dim3(BLOCK), 0, d->stream, B, d->R.p, d->iperm.p, d->Npad);
  launch_check("body place");
}

// host <-> device copies of one body field, host layout [nb][ncomp]
// The whole body record is shared (ArBee.f90 tBody): static properties,
// body-frame and space-frame member offsets and the dynamic state.  Forces
// and torques stay with each system.
void dev_bodies_copy(Device* dst, Device* src) {
  BodyState &a = dst->bodies, &b = src->bodies;
  if (a.nb != b.nb || a.nm != b.nm) fatal("phase space sharing", "rigid bodies do not match");
  if (a.nb == 0) return;
  auto cp = [&](DBuf<double>& x, const DBuf<double>& y) {
    HIP_CHECK(hipMemcpyAsync(x.p, y.p, sizeof(double) * std::min(x.n, y.n), hipMemcpyDeviceToDevice, dst->stream));
  };
  cp(a.mass, b.mass); cp(a.invMass, b.invMass); cp(a.MoI, b.MoI); cp(a.invMoI, b.invMoI); cp(a.coef, b.coef);
  cp(a.d, b.d); cp(a.delta, b.delta);
  cp(a.rcm, b.rcm); cp(a.pcm, b.pcm); cp(a.q, b.q); cp(a.pi, b.pi); cp(a.omega, b.omega);
}

void dev_bodies_get(Device* d, int f, double* host) {
  BodyState& B = d->bodies;
  if (B.nb == 0) return;
  int nc;
  DBuf<double>* a = field(B, f, nc);
  std::vector<double> t((size_t)nc * B.nb);
  HIP_CHECK(hipMemcpyAsync(t.data(), a->p, sizeof(double) * t.size(), hipMemcpyDeviceToHost, d->stream));
  HIP_CHECK(hipStreamSynchronize(d->stream));
  for (int b = 0; b < B.nb; b++) for (int k = 0; k < nc; k++) host[(size_t)nc * b + k] = t[(size_t)k * B.nb + b];
}
void dev_bodies_set(Device* d, int f, const double* host) {
  BodyState& B = d->bodies;
  if (B.nb == 0) return;
  int nc;
  DBuf<double>* a = field(B, f, nc);
  std::vector<double> t((size_t)nc * B.nb);
  for (int b = 0; b < B.nb; b++) for (int k = 0; k < nc; k++) t[(size_t)k * B.nb + b] = host[(size_t)nc * b + k];
  HIP_CHECK(hipMemcpyAsync(a->p, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice, d->stream));
  HIP_CHECK(hipStreamSynchronize(d->stream));
}

void dev_bodies_shadow_pre(Device* d, double* sh, double dt, int mode) {
  BodyState& B = d->bodies;
  if (B.nb == 0) return;
  hipLaunchKernelGGL(k_body_shadow_pre, dim3(div_up(B.nb, BLOCK)), dim3(BLOCK), 0, d->stream, B, sh, dt, mode);
  launch_check("body shadow pre");
}

void dev_bodies_shadow_post(Device* d, const double* sh, double dt, int mode, double* out3) {
  BodyState& B = d->bodies;
  out3[0] = out3[1] = out3[2] = 0.0;
  if (B.nb == 0) return;
  int g = div_up(B.nb, BLOCK);
  hipLaunchKernelGGL(k_body_shadow_post, dim3(g), dim3(BLOCK), 0, d->stream, B, sh, dt, mode);
  launch_check("body shadow post");
  hipLaunchKernelGGL(k_body_reduce<3>, dim3(1), dim3(BLOCK), 0, d->stream, d->result.p + 32, B.part.p, g);
  HIP_CHECK(hipMemcpyAsync(d->hresult + 32, d->result.p + 32, 3 * sizeof(double), hipMemcpyDeviceToHost, d->stream));
  HIP_CHECK(hipStreamSynchronize(d->stream));
  for (int k = 0; k < 3; k++) out3[k] = d->hresult[32 + k];
}

}  // namespace emdee
