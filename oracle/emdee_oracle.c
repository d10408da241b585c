/*
 * emdee_oracle.c -- CPU oracle: FP64 restatement of the reference EmDee engine.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker against which the HIP
 * product library is tested.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load the shared library built from it
 * (oracle/liboracle_emdee.so).  The product path never links, loads or calls
 * it.
 *
 * It exports the same C ABI as include/emdee.h, so one host harness can drive
 * either library.  Every function follows the reference routine cited beside
 * it, in the same arithmetic order (the file:line citations are into the
 * reference source tree, src/).  It is compiled with -ffp-contract=off so that
 * its FP64 results do not depend on FMA contraction.
 *
 * Pinning.  tests/test_oracle_*.py check this restatement against (i) the
 * four [U, W, U+K] energy pins held by the reference's own tests
 * (test/test_pair_lj_*.f90:46; the fourth under the earlier shifted-smoothed
 * rule, see oracle_set_legacy_shifted_smoothing), (ii) brute-force all-pairs neighbor sets,
 * (iii) closed forms of every pair, Coulomb and bonded model, (iv) textbook
 * Ewald results (NaCl Madelung constant) and (v) rigid-body invariants.
 * Functions without such a check are listed as "parity unpinned" in the
 * header comment of the corresponding section below.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "oracle_internal.h"

/* global.f90:31-42 */
static const double PI_ = 3.14159265358979324;
static const double ONE = 1.0, ZERO = 0.0, HALF = 0.5, TWO = 2.0, THIRD = 1.0 / 3.0, FOURTH = 0.25;
#define EPS_ONE 2.220446049250313e-16      /* epsilon(1.0_rb) */

typedef struct { double Potential, Dispersion, Coulomb, Bond, Angle, Dihedral, ShadowPotential; bool UpToDate; } oenergy_t;
typedef struct { double Total, Body; } ovirial_t;

/* ---------------------------------------------------------------- utilities */

/* global.f90:51-56 */
static void oerror(const char *task, const char *msg) {
  fprintf(stderr, "Error in %s: %s.\n", task, msg);
  fflush(stderr);
  exit(1);
}
/* global.f90:60-64 */
static void owarning(const char *msg) { fprintf(stderr, "WARNING: %s.\n", msg); }

static double wtime(void) {
  struct timespec ts; clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s ? s : 1);
  if (!p) oerror("oracle", "out of memory");
  return p;
}

static int thread_num(void) {
#ifdef _OPENMP
  return omp_get_thread_num();
#else
  return 0;
#endif
}

/* Fortran anint: round half away from zero. */
static inline double anint(double x) { return round(x); }
static inline double sign1(double a, double b) { return b >= 0 ? fabs(a) : -fabs(a); }

static int ranged(const int *i, int n, int imax) {
  for (int k = 0; k < n; k++) if (!(i[k] > 0 && i[k] <= imax)) return 0;
  return 1;
}

/* math.f90:695-702 */
static int symm1D(int i, int j) {
  int x = (i < j ? i : j) - 1, y = (i > j ? i : j) - 1;
  return x + (y + 1) * y / 2 + 1;
}

static void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

/* ----------------------------------------------------------- math.f90 bits */

/* A&S 7.1.26: math.f90:35-40 (constants), :685-691 (uerfc) */
static const double A1 = 0.254829592, A2 = -0.284496736, A3 = 1.421413741,
                    A4 = -1.453152027, A5 = 1.061405429, PP = 0.327591100;
static inline double uerfc(double x, double expmx2) {
  double t = ONE / (ONE + PP * x);
  return t * (A1 + t * (A2 + t * (A3 + t * (A4 + t * A5)))) * expmx2;
}

/* math.f90:230-237 */
static double phi_fn(double x) {
  if (fabs(x) > 1e-4) return (ONE - exp(-x)) / x;
  return ONE + HALF * x * (THIRD * x * (ONE - FOURTH * x) - ONE);
}

/* math.f90:642-657 */
static double inverse_of_x_plus_ln_x(double y) {
  double x, x0;
  const double tol = 1.0e-12;
  if (y > 0.5671432904097839) x = y - log(y);
  else x = exp(y);
  x0 = x + ONE;
  while (fabs(x - x0) > tol * x0) {
    x0 = x;
    x = x * (y + ONE - log(x)) / (x + ONE);
  }
  return x;
}

/* KISS generator, math.f90:150-177 */
static int32_t kiss_m(int32_t k, int n) {
  uint32_t u = (uint32_t)k;
  uint32_t s = n > 0 ? (u << n) : (u >> (-n));
  return (int32_t)(u ^ s);
}
static void kiss_init(okiss_t *a, int32_t seed) {
  a->x = kiss_m(kiss_m(kiss_m(seed, 13), -17), 5);
  a->y = kiss_m(kiss_m(kiss_m(a->x, 13), -17), 5);
  a->z = kiss_m(kiss_m(kiss_m(a->y, 13), -17), 5);
  a->w = kiss_m(kiss_m(kiss_m(a->z, 13), -17), 5);
}
static int32_t kiss_i32(okiss_t *a) {
  uint32_t x = (uint32_t)a->x, y = (uint32_t)a->y, z = (uint32_t)a->z, w = (uint32_t)a->w;
  x = 69069u * x + 1327217885u;
  y ^= y << 13; y ^= y >> 17; y ^= y << 5;
  z = 18000u * (z & 65535u) + (z >> 16);
  w = 30903u * (w & 65535u) + (w >> 16);
  a->x = (int32_t)x; a->y = (int32_t)y; a->z = (int32_t)z; a->w = (int32_t)w;
  return (int32_t)(x + y + (z << 16) + w);
}
/* Ziggurat tables (Marsaglia & Tsang 2000): math.f90:80-104 */
static void rng_setup(okiss_t *a, int32_t seed) {
  const double m1 = 2147483648.0;
  double q, dn, tn, vn;
  kiss_init(a, seed);
  a->seeding_required = 0;
  dn = 3.442619855899; tn = 3.442619855899; vn = 0.00991256303526217;
  q = vn * exp(0.5 * dn * dn);
  a->kn[0] = (int32_t)((dn / q) * m1);
  a->kn[1] = 0;
  a->wn[0] = q / m1;
  a->wn[127] = dn / m1;
  a->fn[0] = 1.0;
  a->fn[127] = exp(-0.5 * dn * dn);
  for (int i = 126; i >= 1; i--) {
    dn = sqrt(-2.0 * log(vn / dn + exp(-0.5 * dn * dn)));
    a->kn[i + 1] = (int32_t)((dn / tn) * m1);
    tn = dn;
    a->fn[i] = exp(-0.5 * dn * dn);
    a->wn[i] = dn / m1;
  }
}
static inline int32_t iabs32(int32_t h) { return h == INT32_MIN ? INT32_MIN : (h < 0 ? -h : h); }
/* math.f90:108-146 */
static double rng_normal(okiss_t *a) {
  const double r = 3.442620, s = 0.2328306e-9;
  int32_t hz, iz;
  double x, y;
  hz = kiss_i32(a);
  iz = hz & 127;
  if (iabs32(hz) < a->kn[iz]) return (double)hz * a->wn[iz];
  for (;;) {
    if (iz == 0) {
      for (;;) {
        x = -0.2904764 * log(s * (double)kiss_i32(a) + 0.5);
        y = -log(s * (double)kiss_i32(a) + 0.5);
        if (y + y >= x * x) break;
      }
      double rnor = r + x;
      if (hz <= 0) rnor = -rnor;
      return rnor;
    }
    x = (double)hz * a->wn[iz];
    if (a->fn[iz] + (s * (double)kiss_i32(a) + 0.5) * (a->fn[iz - 1] - a->fn[iz]) < exp(-0.5 * x * x))
      return x;
    hz = kiss_i32(a);
    iz = hz & 127;
    if (iabs32(hz) < a->kn[iz]) return (double)hz * a->wn[iz];
  }
}

/* Shepperd quaternion from a rotation matrix: math.f90:189-218.
   A is column-major A(3,3): A[(r-1) + 3*(c-1)]. */
#define AM(A, r, c) ((A)[((r) - 1) + 3 * ((c) - 1)])
static void quaternion_of(const double *A, double *Q) {
  double Q2[4], a11 = AM(A, 1, 1), a22 = AM(A, 2, 2), a33 = AM(A, 3, 3);
  Q2[0] = ONE + a11 + a22 + a33;
  Q2[1] = ONE + a11 - a22 - a33;
  Q2[2] = ONE - a11 + a22 - a33;
  Q2[3] = ONE - a11 - a22 + a33;
  int imax = 0;
  for (int k = 1; k < 4; k++) if (Q2[k] > Q2[imax]) imax = k;
  double Q2max = Q2[imax], f = sqrt(ONE / Q2max), v[4];
  switch (imax) {
    case 0: v[0] = Q2max; v[1] = AM(A,2,3) - AM(A,3,2); v[2] = AM(A,3,1) - AM(A,1,3); v[3] = AM(A,1,2) - AM(A,2,1); break;
    case 1: v[0] = AM(A,2,3) - AM(A,3,2); v[1] = Q2max; v[2] = AM(A,1,2) + AM(A,2,1); v[3] = AM(A,1,3) + AM(A,3,1); break;
    case 2: v[0] = AM(A,3,1) - AM(A,1,3); v[1] = AM(A,1,2) + AM(A,2,1); v[2] = Q2max; v[3] = AM(A,2,3) + AM(A,3,2); break;
    default: v[0] = AM(A,1,2) - AM(A,2,1); v[1] = AM(A,1,3) + AM(A,3,1); v[2] = AM(A,2,3) + AM(A,3,2); v[3] = Q2max; break;
  }
  for (int k = 0; k < 4; k++) Q[k] = 0.5 * v[k] * f;
}

/* Kopp's analytic 3x3 symmetric eigensolver (Cardano + cross products):
   math.f90:244-424.  matrix/q column-major (3,3). */
static double diag_thresh;
static void compute_eigenvector(double *q, const double *a, double n1tmp, double n2tmp) {
  const double eps = EPS_ONE;
  double norm = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
  double n1 = n1tmp + AM(a, 1, 1) * AM(a, 1, 1);
  double n2 = n2tmp + AM(a, 2, 2) * AM(a, 2, 2);
  double error = n1 * n2, t, f;
  if (n1 <= diag_thresh) { q[0] = ONE; q[1] = ZERO; q[2] = ZERO; }
  else if (n2 <= diag_thresh) { q[0] = ZERO; q[1] = ONE; q[2] = ZERO; }
  else if (norm < (64.0 * eps) * (64.0 * eps) * error) {
    t = fabs(AM(a, 1, 2));
    f = -AM(a, 1, 1) / AM(a, 1, 2);
    if (fabs(AM(a, 2, 2)) > t) { t = fabs(AM(a, 2, 2)); f = -AM(a, 1, 2) / AM(a, 2, 2); }
    if (fabs(AM(a, 2, 3)) > t) f = -AM(a, 1, 3) / AM(a, 2, 3);
    norm = ONE / sqrt(ONE + f * f);
    q[0] = norm; q[1] = f * norm; q[2] = ZERO;
  } else {
    double s = sqrt(ONE / norm);
    q[0] *= s; q[1] *= s; q[2] *= s;
  }
}
static void diagonalization(const double *matrix, double *q, double *w) {
  const double eps = EPS_ONE, sqrt3 = sqrt(3.0);
  double a[9], norm, n1, n2, w1, w2, w3, t, wmax8eps;
  double m, c1, c0, de, dd, ee, ff, p, sqrtp, r, c, s, phi;
  memcpy(a, matrix, sizeof a);
  de = AM(a,1,2) * AM(a,2,3);
  dd = AM(a,1,2) * AM(a,1,2);
  ee = AM(a,2,3) * AM(a,2,3);
  ff = AM(a,1,3) * AM(a,1,3);
  m = AM(a,1,1) + AM(a,2,2) + AM(a,3,3);
  c1 = (AM(a,1,1) * AM(a,2,2) + AM(a,1,1) * AM(a,3,3) + AM(a,2,2) * AM(a,3,3)) - (dd + ee + ff);
  c0 = 27.0 * (AM(a,3,3) * dd + AM(a,1,1) * ee + AM(a,2,2) * ff - AM(a,1,1) * AM(a,2,2) * AM(a,3,3) - TWO * AM(a,1,3) * de);
  p = m * m - 3.0 * c1;
  r = m * (p - 1.5 * c1) - HALF * c0;
  sqrtp = sqrt(fabs(p));
  phi = THIRD * atan2(sqrt(fabs(6.75 * c1 * c1 * (p - c1) + c0 * (r + 0.25 * c0))), r);
  c = sqrtp * cos(phi);
  s = (ONE / sqrt3) * sqrtp * sin(phi);
  p = THIRD * (m - c);
  w1 = p + c; w2 = p - s; w3 = p + s;
#define SWAPD(u, v) do { double _t = u; u = v; v = _t; } while (0)
  if (fabs(w1) < fabs(w3)) SWAPD(w1, w3);
  if (fabs(w1) < fabs(w2)) SWAPD(w1, w2);
  if (fabs(w2) < fabs(w3)) SWAPD(w2, w3);
  w[0] = w1; w[1] = w2; w[2] = w3;
  wmax8eps = 8.0 * eps * fabs(w1);
  diag_thresh = wmax8eps * wmax8eps;
  n1 = AM(a,1,2) * AM(a,1,2) + AM(a,1,3) * AM(a,1,3);
  n2 = AM(a,1,2) * AM(a,1,2) + AM(a,2,3) * AM(a,2,3);
  AM(q,1,1) = AM(a,1,2) * AM(a,2,3) - AM(a,1,3) * AM(a,2,2);
  AM(q,1,2) = AM(q,1,1);
  AM(q,2,1) = AM(a,1,3) * AM(a,1,2) - AM(a,2,3) * AM(a,1,1);
  AM(q,2,2) = AM(q,2,1);
  AM(q,3,2) = AM(a,1,2) * AM(a,1,2);
  AM(a,1,1) = AM(a,1,1) - w1;
  AM(a,2,2) = AM(a,2,2) - w1;
  {
    double v[3] = { AM(q,1,2) + AM(a,1,3) * w1, AM(q,2,2) + AM(a,2,3) * w1, AM(a,1,1) * AM(a,2,2) - AM(q,3,2) };
    compute_eigenvector(v, a, n1, n2);
    AM(q,1,1) = v[0]; AM(q,2,1) = v[1]; AM(q,3,1) = v[2];
  }
  t = w1 - w2;
  if (fabs(t) > wmax8eps) {
    AM(a,1,1) = AM(a,1,1) + t;
    AM(a,2,2) = AM(a,2,2) + t;
    double v[3] = { AM(q,1,2) + AM(a,1,3) * w2, AM(q,2,2) + AM(a,2,3) * w2, AM(a,1,1) * AM(a,2,2) - AM(q,3,2) };
    compute_eigenvector(v, a, n1, n2);
    AM(q,1,2) = v[0]; AM(q,2,2) = v[1]; AM(q,3,2) = v[2];
  } else {
    AM(a,2,1) = AM(a,1,2);
    AM(a,3,1) = AM(a,1,3);
    AM(a,3,2) = AM(a,2,3);
    AM(a,1,1) = AM(a,1,1) + w1;
    AM(a,2,2) = AM(a,2,2) + w1;
    int i = 0, success = 0;
    while (i < 3 && !success) {
      i = i + 1;
      AM(a,i,i) = AM(a,i,i) - w2;
      n1 = AM(a,1,i) * AM(a,1,i) + AM(a,2,i) * AM(a,2,i) + AM(a,3,i) * AM(a,3,i);
      success = n1 > diag_thresh;
      if (success) {
        double q1[3] = { AM(q,1,1), AM(q,2,1), AM(q,3,1) }, ai[3] = { AM(a,1,i), AM(a,2,i), AM(a,3,i) }, v[3];
        cross3(q1, ai, v);
        norm = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        success = norm > (256.0 * eps) * (256.0 * eps) * n1;
        if (success) { double f = sqrt(ONE / norm); v[0] *= f; v[1] *= f; v[2] *= f; }
        AM(q,1,2) = v[0]; AM(q,2,2) = v[1]; AM(q,3,2) = v[2];
      }
    }
    if (!success) {
      int ii = 1, j;
      while (AM(q, ii, 1) == ZERO) ii = ii + 1;
      j = 1 + ii % 3;
      norm = ONE / sqrt(AM(q,ii,1) * AM(q,ii,1) + AM(q,j,1) * AM(q,j,1));
      AM(q,ii,2) = AM(q,j,1) * norm;
      AM(q,j,2) = -AM(q,ii,1) * norm;
      AM(q, 1 + (ii + 1) % 3, 2) = ZERO;
    }
  }
  {
    double q1[3] = { AM(q,1,1), AM(q,2,1), AM(q,3,1) }, q2[3] = { AM(q,1,2), AM(q,2,2), AM(q,3,2) }, v[3];
    cross3(q1, q2, v);
    AM(q,1,3) = v[0]; AM(q,2,3) = v[1]; AM(q,3,3) = v[2];
  }
}

/* math.f90:428-436 */
static int staircase(double x) {
  if (x > ZERO) return (int)ceil(x - HALF);
  return (int)floor(x + HALF);
}

/* Jacobi elliptic functions by the AGM / descending Landen method:
   math.f90:440-501 */
static void jacobi(double u, double m, double *jac) {
  enum { NN = 16 };
  int n;
  double mu[NN], nu[NN], c[NN], d[NN];
  double sin_umu, cos_umu, t, r, sn, cn, dn;
  if (fabs(m) > ONE) { jac[0] = jac[1] = jac[2] = NAN; return; }
  if (fabs(m) < TWO * EPS_ONE) { jac[0] = sin(u); jac[1] = cos(u); jac[2] = ONE; return; }
  if (fabs(m - ONE) < TWO * EPS_ONE) { jac[0] = tanh(u); jac[1] = jac[2] = ONE / cosh(u); return; }
  n = 0;
  mu[0] = ONE;
  nu[0] = sqrt(ONE - m);
  while (fabs(mu[n] - nu[n]) > 4.0 * EPS_ONE * fabs(mu[n] + nu[n])) {
    mu[n + 1] = HALF * (mu[n] + nu[n]);
    nu[n + 1] = sqrt(mu[n] * nu[n]);
    n = n + 1;
    if (n >= NN - 1) { jac[0] = jac[1] = jac[2] = NAN; return; }
  }
  sin_umu = sin(u * mu[n]);
  cos_umu = cos(u * mu[n]);
  if (fabs(sin_umu) < fabs(cos_umu)) {
    t = sin_umu / cos_umu;
    c[n] = mu[n] * t;
    d[n] = ONE;
    while (n > 0) {
      c[n - 1] = d[n] * c[n];
      r = (c[n] * c[n]) / mu[n];
      n = n - 1;
      d[n] = (r + nu[n]) / (r + mu[n]);
    }
    dn = sqrt(ONE - m) / d[n];
    cn = dn * sign1(ONE, cos_umu) / hypot(ONE, c[n]);
    sn = cn * c[n] / sqrt(ONE - m);
  } else {
    t = cos_umu / sin_umu;
    c[n] = mu[n] * t;
    d[n] = ONE;
    while (n > 0) {
      c[n - 1] = d[n] * c[n];
      r = (c[n] * c[n]) / mu[n];
      n = n - 1;
      d[n] = (r + nu[n]) / (r + mu[n]);
    }
    dn = d[n];
    sn = sign1(ONE, sin_umu) / hypot(ONE, c[n]);
    cn = c[n] * sn;
  }
  jac[0] = sn; jac[1] = cn; jac[2] = dn;
}

/* Carlson symmetric integrals (Carlson 1995 duplication algorithm):
   math.f90:505-638 */
#define TINY_D 2.2250738585072014e-308
#define HUGE_D 1.7976931348623157e308
static double Carlson_RF(double x, double y, double z) {
  const double errtol = 0.001, c1 = 1.0 / 24.0, c2 = 3.0 / 44.0, c3 = 1.0 / 14.0;
  const double lolim = cbrt(5.0 * TINY_D), uplim = 0.3 * cbrt(0.2 * HUGE_D);
  double e2, e3, lamda, mu, s, xn, xndev, xnroot, yn, yndev, ynroot, zn, zndev, znroot, eps;
  double mn = fmin(x, fmin(y, z)), mx = fmax(x, fmax(y, z));
  double ms = fmin(x + y, fmin(x + z, y + z));
  if (mn < ZERO || mx > uplim || ms < lolim) return NAN;
  xn = x; yn = y; zn = z;
  for (;;) {
    mu = (xn + yn + zn) / 3.0;
    xndev = 2.0 - (mu + xn) / mu;
    yndev = 2.0 - (mu + yn) / mu;
    zndev = 2.0 - (mu + zn) / mu;
    eps = fmax(fabs(xndev), fmax(fabs(yndev), fabs(zndev)));
    if (eps < errtol) break;
    xnroot = sqrt(xn); ynroot = sqrt(yn); znroot = sqrt(zn);
    lamda = xnroot * (ynroot + znroot) + ynroot * znroot;
    xn = (xn + lamda) * 0.250;
    yn = (yn + lamda) * 0.250;
    zn = (zn + lamda) * 0.250;
  }
  e2 = xndev * yndev - zndev * zndev;
  e3 = xndev * yndev * zndev;
  s = 1.0 + (c1 * e2 - 0.10 - c2 * e3) * e2 + c3 * e3;
  return s / sqrt(mu);
}
static double Carlson_RC(double x, double y) {
  const double errtol = 0.001, c1 = 1.0 / 7.0, c2 = 9.0 / 22.0;
  const double lolim = 5.0 * TINY_D, uplim = 0.2 * HUGE_D;
  double lamda, mu, s, sn, xn, yn;
  if (x < ZERO || y <= ZERO || fmax(x, y) > uplim || x + y < lolim) return NAN;
  xn = x; yn = y;
  for (;;) {
    mu = (xn + yn + yn) / 3.0;
    sn = (yn + mu) / mu - 2.0;
    if (fabs(sn) < errtol) break;
    lamda = 2.0 * sqrt(xn) * sqrt(yn) + yn;
    xn = (xn + lamda) * 0.250;
    yn = (yn + lamda) * 0.250;
  }
  s = sn * sn * (0.30 + sn * (c1 + sn * (0.3750 + sn * c2)));
  return (1.0 + s) / sqrt(mu);
}
static double Carlson_RJ(double x, double y, double z, double p) {
  const double errtol = 0.001, c1 = 3.0 / 14.0, c2 = 1.0 / 3.0, c3 = 3.0 / 22.0, c4 = 3.0 / 26.0;
  const double lolim = cbrt(5.0 * TINY_D), uplim = 0.30 * cbrt(0.2 * HUGE_D);
  double alfa, beta, ea, eb, ec, e2, e3, eps, lamda, mu, pn, pndev, power4, sigma, s1, s2, s3;
  double xn, xndev, xnroot, yn, yndev, ynroot, zn, zndev, znroot;
  double mn = fmin(x, fmin(y, z)), mx = fmax(fmax(x, y), fmax(z, p));
  double ms = fmin(fmin(x + y, x + z), fmin(y + z, p));
  if (mn < ZERO || mx > uplim || ms < lolim) return NAN;
  xn = x; yn = y; zn = z; pn = p;
  sigma = 0.0; power4 = 1.0;
  for (;;) {
    mu = (xn + yn + zn + pn + pn) * 0.20;
    xndev = (mu - xn) / mu;
    yndev = (mu - yn) / mu;
    zndev = (mu - zn) / mu;
    pndev = (mu - pn) / mu;
    eps = fmax(fmax(fabs(xndev), fabs(yndev)), fmax(fabs(zndev), fabs(pndev)));
    if (eps < errtol) break;
    xnroot = sqrt(xn); ynroot = sqrt(yn); znroot = sqrt(zn);
    lamda = xnroot * (ynroot + znroot) + ynroot * znroot;
    alfa = pn * (xnroot + ynroot + znroot) + xnroot * ynroot * znroot;
    alfa = alfa * alfa;
    beta = pn * (pn + lamda) * (pn + lamda);
    sigma = sigma + power4 * Carlson_RC(alfa, beta);
    power4 = power4 * 0.250;
    xn = (xn + lamda) * 0.250;
    yn = (yn + lamda) * 0.250;
    zn = (zn + lamda) * 0.250;
    pn = (pn + lamda) * 0.250;
  }
  ea = xndev * (yndev + zndev) + yndev * zndev;
  eb = xndev * yndev * zndev;
  ec = pndev * pndev;
  e2 = ea - 3.0 * ec;
  e3 = eb + 2.0 * pndev * (ea - ec);
  s1 = 1.0 + e2 * (-c1 + 0.750 * c3 * e2 - 1.50 * c4 * e3);
  s2 = eb * (0.50 * c2 + pndev * (-c3 - c3 + pndev * c4));
  s3 = pndev * ea * (c2 - pndev * c3) - c2 * pndev * ec;
  return 3.0 * sigma + power4 * (s1 + s2 + s3) / (mu * sqrt(mu));
}

/* Exposed for tests: special functions that have textbook checks. */
double oracle_uerfc(double x) { return uerfc(x, exp(-x * x)); }
double oracle_carlson_rf(double x, double y, double z) { return Carlson_RF(x, y, z); }
double oracle_carlson_rj(double x, double y, double z, double p) { return Carlson_RJ(x, y, z, p); }
void oracle_jacobi(double u, double m, double *jac) { jacobi(u, m, jac); }
void oracle_diagonalization(const double *matrix, double *q, double *w) { diagonalization(matrix, q, w); }
double oracle_inverse_x_plus_ln_x(double y) { return inverse_of_x_plus_ln_x(y); }
double oracle_phi(double x) { return phi_fn(x); }
/* Counter-based generator of the batched hybrid Monte Carlo driver (no
   reference counterpart; used by the tests that replay batched HMC runs on
   the reference API).  splitmix64's output function (Steele, Lea & Flood,
   "Fast splittable pseudorandom number generators", OOPSLA 2014) keyed by
   (seed, stream, counter); Gaussians by the Box-Muller transform. */
static uint64_t sm64_mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
/* the classic sequential generator: state += golden gamma, output mix */
uint64_t oracle_splitmix64_next(uint64_t *state) {
  *state += 0x9E3779B97F4A7C15ULL;
  return sm64_mix(*state);
}
double oracle_hmc_uniform(uint64_t seed, uint64_t stream, uint64_t counter) {
  uint64_t x = sm64_mix(seed ^ sm64_mix(stream * 0x9E3779B97F4A7C15ULL + counter));
  return ((double)(x >> 11) + 0.5) / 9007199254740992.0;
}
static uint64_t hmc_stream_of(int replica, int cycle) { return ((uint64_t)(uint32_t)replica << 32) | (uint32_t)cycle; }
void oracle_hmc_momenta(uint64_t seed, int replica, int cycle, int n, double kT, double *P) {
  uint64_t st = hmc_stream_of(replica, cycle);
  double sq = sqrt(kT);
  for (int i = 0; i < n; i++)
    for (int x = 0; x < 3; x++) {
      double u1 = oracle_hmc_uniform(seed, st, 6ULL * i + 2 * x), u2 = oracle_hmc_uniform(seed, st, 6ULL * i + 2 * x + 1);
      P[3 * i + x] = sq * (sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
    }
}
double oracle_hmc_accept_uniform(uint64_t seed, int replica, int cycle) {
  return oracle_hmc_uniform(seed, hmc_stream_of(replica, cycle), 1ULL << 40);
}

/* First n normals of a freshly seeded generator. */
void oracle_normals(int seed, int n, double *out) {
  okiss_t g; rng_setup(&g, seed);
  for (int i = 0; i < n; i++) out[i] = rng_normal(&g);
}
/* First n raw KISS integers. */
void oracle_kiss(int seed, int n, int *out) {
  okiss_t g; kiss_init(&g, seed);
  for (int i = 0; i < n; i++) out[i] = kiss_i32(&g);
}

/* ================================================================== models */

static omodel_t *new_model(int kind) {
  omodel_t *m = xcalloc(1, sizeof *m);
  m->kind = kind;
  return m;
}

static int is_pair(const omodel_t *m) { return m->kind >= K_PAIR_NONE && m->kind <= K_PAIR_SOFTCORE_CUT; }
static int is_coul(const omodel_t *m) { return m->kind >= K_COUL_NONE && m->kind <= K_COUL_SHIFTED_SQUARE_SMOOTHED; }
static int is_nonbonded(const omodel_t *m) { return is_pair(m) || is_coul(m); }

/* pair_lj_cut.f90:52-69 */
static void lj_setup(omodel_t *m, double eps, double sig) {
  m->epsilon = eps; m->sigma = sig;
  m->eps4 = 4.0 * m->epsilon;
  m->eps24 = 24.0 * m->epsilon;
  m->sigsq = m->sigma * m->sigma;
}
/* pair_softcore_cut.f90:58-82 */
static void softcore_setup(omodel_t *m, double eps, double sig, double lam) {
  m->epsilon = eps; m->sigma = sig; m->lambda = lam;
  if (m->lambda < ZERO || m->lambda > ONE) oerror("pair_softcore_cut setup", "out-of-range parameter lambda");
  m->prefactor = 4.0 * m->epsilon * m->lambda;         /* lambda**exponent_n, n = 1 */
  m->prefactor6 = 6.0 * m->prefactor;
  m->invSigSq = ONE / (m->sigma * m->sigma);
  m->shift = HALF * (ONE - m->lambda);                 /* alpha*(1-lambda)**p, p = 1 */
}

void *EmDee_pair_none(void) { return new_model(K_PAIR_NONE); }
void *EmDee_coul_none(void) { return new_model(K_COUL_NONE); }
void *EmDee_bond_none(void) { return new_model(K_BOND_NONE); }
void *EmDee_angle_none(void) { return new_model(K_ANGLE_NONE); }
void *EmDee_dihedral_none(void) { return new_model(K_DIHEDRAL_NONE); }

void *EmDee_pair_lj_cut(double epsilon, double sigma) {
  omodel_t *m = new_model(K_PAIR_LJ_CUT); lj_setup(m, epsilon, sigma); return m;
}
void *EmDee_pair_softcore_cut(double epsilon, double sigma, double lambda) {
  omodel_t *m = new_model(K_PAIR_SOFTCORE_CUT); softcore_setup(m, epsilon, sigma, lambda); return m;
}
void *EmDee_coul_cut(void) { return new_model(K_COUL_CUT); }
void *EmDee_coul_sf(void) { omodel_t *m = new_model(K_COUL_SF); m->shifted_force = 1; return m; }   /* coul_sf.f90:52 */
void *EmDee_coul_damped(double damp) {                                                               /* coul_damped.f90:56-66 */
  omodel_t *m = new_model(K_COUL_DAMPED);
  m->damp = damp; m->alpha = damp; m->beta = TWO * m->alpha / sqrt(PI_); return m;
}
void *EmDee_coul_long(void) { omodel_t *m = new_model(K_COUL_LONG); m->requires_kspace = 1; return m; }
void *EmDee_coul_damped_smoothed(double damp, double skinWidth) {                                    /* coul_damped_smoothed.f90:55-71 */
  omodel_t *m = new_model(K_COUL_DAMPED_SMOOTHED);
  m->damp = damp; m->skinWidth = skinWidth; m->alpha = m->damp; m->beta = TWO * m->alpha / sqrt(PI_); return m;
}
void *EmDee_coul_damped_square_smoothed(double damp, double skinWidth) {
  omodel_t *m = new_model(K_COUL_DAMPED_SQUARE_SMOOTHED);
  m->damp = damp; m->skinWidth = skinWidth; m->alpha = m->damp; m->beta = TWO * m->alpha / sqrt(PI_); return m;
}
void *EmDee_coul_square_smoothed(double skinWidth) {
  omodel_t *m = new_model(K_COUL_SQUARE_SMOOTHED); m->skinWidth = skinWidth; return m;
}
void *EmDee_coul_shifted_square_smoothed(double skinWidth) {
  omodel_t *m = new_model(K_COUL_SHIFTED_SQUARE_SMOOTHED); m->skinWidth = skinWidth; m->shifted = 1; return m;
}
/* bond_harmonic.f90:50-65, angle_harmonic.f90:50-65 */
void *EmDee_bond_harmonic(double k, double r0) {
  omodel_t *m = new_model(K_BOND_HARMONIC);
  m->k = k; m->r0 = r0; m->minus_k = -m->k; m->half_k = HALF * m->k; return m;
}
void *EmDee_angle_harmonic(double k, double theta0) {
  omodel_t *m = new_model(K_ANGLE_HARMONIC);
  m->k = k; m->r0 = theta0; m->minus_k = -m->k; m->half_k = HALF * m->k; return m;
}
/* A harmonic dihedral, E = k/2 (phi - phi0)^2 with the difference taken in
   (-pi, pi]; the reference defines the dihedral model class
   (modelClass_dihedral.f90:28-45: compute(Ed, Fd, phi), factor14) but ships
   no concrete model.  Named and laid out after bond_harmonic and
   angle_harmonic; factor14 = 0 (no scaled 1-4 pair term). */
void *EmDee_dihedral_harmonic(double k, double phi0) {
  omodel_t *m = new_model(K_DIHEDRAL_HARMONIC);
  m->k = k; m->r0 = phi0; m->minus_k = -m->k; m->half_k = HALF * m->k; m->factor14 = ZERO; return m;
}
void *EmDee_kspace_ewald(double accuracy) {                                                          /* kspace_ewald.f90:63-78 */
  omodel_t *m = new_model(K_KSPACE_EWALD); m->accuracy = accuracy; return m;
}

/* Modifier constructors: modelClass_nonbonded.f90:83-241 */
static void *modified(void *model, int modifier, double skin, int set_skin, const char *task) {
  if (!model) oerror(task, "a valid pair model must be provided");
  omodel_t *src = model;
  if (!is_nonbonded(src)) oerror(task, "a valid pair model must be provided");
  omodel_t *m = xcalloc(1, sizeof *m);
  *m = *src;
  m->modifier = modifier;
  if (set_skin) m->skin = skin;
  return m;
}
void *EmDee_shifted(void *model) { return modified(model, MOD_SHIFTED, 0, 0, "shifted potential assignment"); }
void *EmDee_shifted_force(void *model) { return modified(model, MOD_SHIFTED_FORCE, 0, 0, "shifted-force potential assignment"); }
void *EmDee_smoothed(void *model, double skin) { return modified(model, MOD_SMOOTHED, skin, 1, "smoothed potential assignment"); }
void *EmDee_shifted_smoothed(void *model, double skin) { return modified(model, MOD_SHIFTED_SMOOTHED, skin, 1, "shifted-smoothed potential assignment"); }
void *EmDee_square_smoothed(void *model, double skin) { return modified(model, MOD_SQUARE_SMOOTHED, skin, 1, "square-smoothed potential assignment"); }
void *EmDee_shifted_square_smoothed(void *model, double skin) { return modified(model, MOD_SHIFTED_SQUARE_SMOOTHED, skin, 1, "shifted-square-smoothed potential assignment"); }

/* ---------------------------------------------- pair model compute/energy */

/* pair_lj_cut.f90:73-86, pair_softcore_cut.f90:86-101 */
static void pair_compute(const omodel_t *m, double *E, double *W, double invR, double invR2) {
  (void)invR;
  switch (m->kind) {
    case K_PAIR_LJ_CUT: {
      double sr2 = m->sigsq * invR2, sr6 = sr2 * sr2 * sr2, sr12 = sr6 * sr6;
      *E = m->eps4 * (sr12 - sr6);
      *W = m->eps24 * (sr12 + sr12 - sr6);
      break;
    }
    case K_PAIR_SOFTCORE_CUT: {
      double rsig2 = m->invSigSq / invR2, rsig6 = rsig2 * rsig2 * rsig2;
      double sinv = ONE / (rsig6 + m->shift), sinvSq = sinv * sinv, sinvCb = sinv * sinvSq;
      *E = m->prefactor * (sinvSq - sinv);
      *W = m->prefactor6 * rsig6 * (sinvCb + sinvCb - sinvSq);
      break;
    }
    default: *E = ZERO; *W = ZERO;
  }
}
static double pair_energy(const omodel_t *m, double invR, double invR2) {
  (void)invR;
  switch (m->kind) {
    case K_PAIR_LJ_CUT: {
      double sr2 = m->sigsq * invR2, sr6 = sr2 * sr2 * sr2, sr12 = sr6 * sr6;
      return m->eps4 * (sr12 - sr6);
    }
    case K_PAIR_SOFTCORE_CUT: {
      double rsig2 = m->invSigSq / invR2, rsig6 = rsig2 * rsig2 * rsig2;
      double sinv = ONE / (rsig6 + m->shift);
      return m->prefactor * (sinv * sinv - sinv);
    }
    default: return ZERO;
  }
}

/* ------------------------------------------- coulomb model compute/energy */

static inline double smooth_G(double u) { double u2 = u * u, u3 = u * u2; return 1.0 + u3 * (15.0 * u - 6.0 * u2 - 10.0); }

/* coul_cut.f90:62-70, coul_sf.f90:61-72, coul_damped.f90:70-83, coul_long.f90:80-92,
   coul_damped_smoothed.f90:92-115, coul_damped_square_smoothed.f90:92-115,
   coul_square_smoothed.f90:85-104, coul_shifted_square_smoothed.f90:88-107 */
static void coul_compute(const omodel_t *m, double *E, double *W, double invR, double invR2) {
  switch (m->kind) {
    case K_COUL_CUT: *E = invR; *W = *E; break;
    case K_COUL_SF: {
      double rFc = m->fshift / invR;
      *E = invR + m->eshift + rFc;
      *W = invR - rFc;
      break;
    }
    case K_COUL_DAMPED: case K_COUL_LONG: {
      double x = m->alpha / invR, expmx2 = exp(-x * x);
      *E = uerfc(x, expmx2) * invR;
      *W = *E + m->beta * expmx2;
      break;
    }
    case K_COUL_DAMPED_SMOOTHED: {
      double r = ONE / invR, x = m->alpha * r, expmx2 = exp(-x * x);
      double Eij = uerfc(x, expmx2) * invR, Wij = Eij + m->beta * expmx2;
      if (r > m->Rm) {
        double u = m->factor * (r - m->Rm), u2 = u * u, u3 = u * u2;
        double G = 1.0 + u3 * (15.0 * u - 6.0 * u2 - 10.0);
        double WG = -30.0 * u2 * (2.0 * u - u2 - 1.0) * m->factor * r;
        Wij = Wij * G + Eij * WG;
        Eij = Eij * G;
      }
      *E = Eij; *W = Wij;
      break;
    }
    case K_COUL_DAMPED_SQUARE_SMOOTHED: {
      double x = m->alpha / invR, expmx2 = exp(-x * x);
      double Eij = uerfc(x, expmx2) * invR, Wij = Eij + m->beta * expmx2;
      if (invR < m->invRm) {
        double r2 = ONE / invR2, u = m->factor * (r2 - m->Rm2), u2 = u * u, u3 = u * u2;
        double G = 1.0 + u3 * (15.0 * u - 6.0 * u2 - 10.0);
        double WG = -60.0 * u2 * (2.0 * u - u2 - 1.0) * m->factor * r2;
        Wij = Wij * G + Eij * WG;
        Eij = Eij * G;
      }
      *E = Eij; *W = Wij;
      break;
    }
    case K_COUL_SQUARE_SMOOTHED: case K_COUL_SHIFTED_SQUARE_SMOOTHED: {
      double Wij = invR;
      double Eij = (m->kind == K_COUL_SQUARE_SMOOTHED) ? Wij : Wij + m->eshift;
      if (invR < m->invRm) {
        double r2 = ONE / invR2, u = m->factor * (r2 - m->Rm2), u2 = u * u, u3 = u * u2;
        double G = 1.0 + u3 * (15.0 * u - 6.0 * u2 - 10.0);
        double WG = -60.0 * u2 * (2.0 * u - u2 - 1.0) * m->factor * r2;
        Wij = Wij * G + Eij * WG;
        Eij = Eij * G;
      }
      *E = Eij; *W = Wij;
      break;
    }
    default: *E = ZERO; *W = ZERO;
  }
}
static double coul_energy(const omodel_t *m, double invR, double invR2) {
  double E, W;
  switch (m->kind) {
    case K_COUL_SF: return invR + m->eshift + m->fshift / invR;
    case K_COUL_DAMPED: case K_COUL_LONG: { double x = m->alpha / invR; return uerfc(x, exp(-x * x)) * invR; }
    case K_COUL_DAMPED_SMOOTHED: {
      double r = ONE / invR, x = m->alpha * r, Eij = uerfc(x, exp(-x * x)) * invR;
      if (r > m->Rm) Eij = Eij * smooth_G(m->factor * (r - m->Rm));
      return Eij;
    }
    case K_COUL_DAMPED_SQUARE_SMOOTHED: {
      double x = m->alpha / invR, Eij = uerfc(x, exp(-x * x)) * invR;
      if (invR < m->invRm) Eij = Eij * smooth_G(m->factor * (ONE / invR2 - m->Rm2));
      return Eij;
    }
    case K_COUL_SQUARE_SMOOTHED: case K_COUL_SHIFTED_SQUARE_SMOOTHED: {
      double Eij = (m->kind == K_COUL_SQUARE_SMOOTHED) ? invR : invR + m->eshift;
      if (invR < m->invRm) Eij = Eij * smooth_G(m->factor * (ONE / invR2 - m->Rm2));
      return Eij;
    }
    default: coul_compute(m, &E, &W, invR, invR2); return E;
  }
}
/* Virial-only variants: they differ from "compute" in the square-smoothed
   Coulomb models (coul_square_smoothed.f90:131-148,
   coul_shifted_square_smoothed.f90:134-151). */
static double coul_virial(const omodel_t *m, double invR, double invR2) {
  double E, W;
  switch (m->kind) {
    case K_COUL_SQUARE_SMOOTHED: {
      double Wij = invR;
      if (invR < m->invRm) {
        double r2 = ONE / invR2, u = m->factor * (r2 - m->Rm2), u2 = u * u, u3 = u * u2;
        double G = 1.0 + u3 * (15.0 * u - 6.0 * u2 - 10.0);
        double WG = -60.0 * u2 * (2.0 * u - u2 - 1.0) * m->factor * r2;
        Wij = Wij * (G + WG);
      }
      return Wij;
    }
    case K_COUL_SHIFTED_SQUARE_SMOOTHED: {
      double Wij = invR;
      if (invR < m->invRm) {
        double r2 = ONE / invR2, u = m->factor * (r2 - m->Rm2), u2 = u * u, u3 = u * u2;
        double G = 1.0 + u3 * (15.0 * u - 6.0 * u2 - 10.0);
        double WG = -60.0 * u2 * (2.0 * u - u2 - 1.0) * m->factor * r2;
        Wij = Wij * G + (Wij + m->eshift) * WG;
      }
      return Wij;
    }
    default: coul_compute(m, &E, &W, invR, invR2); return W;
  }
}

static void nb_compute(const omodel_t *m, double *E, double *W, double invR, double invR2) {
  if (is_pair(m)) pair_compute(m, E, W, invR, invR2); else coul_compute(m, E, W, invR, invR2);
}
static double nb_energy(const omodel_t *m, double invR, double invR2) {
  return is_pair(m) ? pair_energy(m, invR, invR2) : coul_energy(m, invR, invR2);
}

/* Earlier shifted-smoothed rule, kept only so that the checker can reproduce
   the energy pin of test/test_pair_lj_shifted_smoothed.f90:46, which was
   produced by an earlier version of the library than the src/ snapshot: the
   energy shift is -E(rc) (not the snapshot's -(E(Rm) + E(rc))/2,
   modelClass_nonbonded.f90:281-290) and it is added to the pair energy only
   inside the switching region (the snapshot adds it at every distance,
   apply_modifier.f90:17-20,38-41).  Off unless a test turns it on. */
static int legacy_shifted_smoothing = 0;
void oracle_set_legacy_shifted_smoothing(int on) { legacy_shifted_smoothing = on; }

/* modelClass_nonbonded.f90:245-296 */
static void modifier_setup(omodel_t *m, double cutoff) {
  double Ec = 0, Wc = 0, Es, Ws;
  m->fshift = ZERO;
  m->eshift = ZERO;
  m->Rm = cutoff - m->skin;
  m->RmSq = m->Rm * m->Rm;
  int shifting = m->modifier == MOD_SHIFTED || m->modifier == MOD_SHIFTED_FORCE ||
                 m->modifier == MOD_SHIFTED_SMOOTHED || m->modifier == MOD_SHIFTED_SQUARE_SMOOTHED;
  if (shifting) nb_compute(m, &Ec, &Wc, ONE / cutoff, ONE / (cutoff * cutoff));
  switch (m->modifier) {
    case MOD_SHIFTED: m->eshift = -Ec; break;
    case MOD_SHIFTED_FORCE: m->eshift = -(Ec + Wc); m->fshift = Wc / cutoff; break;
    case MOD_SMOOTHED: case MOD_SHIFTED_SMOOTHED:
      if (shifting) { nb_compute(m, &Es, &Ws, ONE / m->Rm, ONE / m->RmSq); m->eshift = legacy_shifted_smoothing ? -Ec : -HALF * (Es + Ec); }
      m->factor = ONE / (cutoff - m->Rm);
      m->Rm2fac = m->factor * m->Rm;
      break;
    case MOD_SQUARE_SMOOTHED: case MOD_SHIFTED_SQUARE_SMOOTHED:
      if (shifting) { nb_compute(m, &Es, &Ws, ONE / m->Rm, ONE / m->RmSq); m->eshift = legacy_shifted_smoothing ? -Ec : -HALF * (Es + Ec); }
      m->factor = ONE / (cutoff * cutoff - m->RmSq);
      m->Rm2fac = m->factor * m->RmSq;
      break;
    default: break;
  }
}

/* coul_*.f90 apply_cutoff: coul_damped_smoothed.f90:76-85 and the
   square-smoothed variants. */
static void coul_apply_cutoff(omodel_t *m, double Rc) {
  switch (m->kind) {
    case K_COUL_DAMPED_SMOOTHED:
      m->Rm = Rc - m->skinWidth; m->Rm2 = m->Rm * m->Rm; m->invRm = ONE / m->Rm; m->factor = ONE / (Rc - m->Rm);
      break;
    case K_COUL_DAMPED_SQUARE_SMOOTHED: case K_COUL_SQUARE_SMOOTHED: case K_COUL_SHIFTED_SQUARE_SMOOTHED:
      m->Rm2 = (Rc - m->skinWidth) * (Rc - m->skinWidth);
      m->invRm = ONE / (Rc - m->skinWidth);
      m->factor = ONE / (Rc * Rc - m->Rm2);
      break;
    default: break;
  }
}
/* modelClass_coul.f90:62-93 */
static void coul_cutoff_setup(omodel_t *m, double cutoff) {
  double invR, invR2, E, W;
  m->fshift = ZERO;
  m->eshift = ZERO;
  if (m->shifted || m->shifted_force) {
    invR = ONE / cutoff;
    invR2 = invR * invR;
    coul_compute(m, &E, &W, invR, invR2);
    if (m->shifted_force) { m->fshift = W / cutoff; m->eshift = -(E + W); }
    else { m->fshift = ZERO; m->eshift = -E; }
  }
  coul_apply_cutoff(m, cutoff);
}
/* coul_long.f90:66-73, modelClass_coul.f90:106-111 */
static void coul_kspace_setup(omodel_t *m, double alpha) {
  if (m->kind == K_COUL_LONG) { m->alpha = alpha; m->beta = TWO * alpha / sqrt(PI_); }
  else if (m->requires_kspace) m->alpha = alpha;
}

static const char *model_name(const omodel_t *m) {
  switch (m->kind) {
    case K_PAIR_LJ_CUT: return "lj_cut";
    case K_PAIR_SOFTCORE_CUT: return "softcore_cut";
    default: return "none";
  }
}

/* Mixing rules: pair_lj_cut.f90:122-138, pair_softcore_cut.f90:139-163,
   modelClass_pair.f90:190-200 (none) -- "this%mix(other)". Returns 0 when no
   rule applies. */
static int pair_mix_rule(const omodel_t *thiz, const omodel_t *other, omodel_t *out) {
  memset(out, 0, sizeof *out);
  switch (thiz->kind) {
    case K_PAIR_NONE: out->kind = K_PAIR_NONE; return 1;
    case K_PAIR_LJ_CUT:
      if (other->kind == K_PAIR_LJ_CUT) {
        out->kind = K_PAIR_LJ_CUT;
        lj_setup(out, sqrt(thiz->epsilon * other->epsilon), HALF * (thiz->sigma + other->sigma));
        return 1;
      }
      return 0;
    case K_PAIR_SOFTCORE_CUT:
      if (other->kind == K_PAIR_SOFTCORE_CUT) {
        out->kind = K_PAIR_SOFTCORE_CUT;
        softcore_setup(out, sqrt(thiz->epsilon * other->epsilon), HALF * (thiz->sigma + other->sigma), thiz->lambda * other->lambda);
        return 1;
      }
      if (other->kind == K_PAIR_LJ_CUT) {
        out->kind = K_PAIR_SOFTCORE_CUT;
        softcore_setup(out, sqrt(thiz->epsilon * other->epsilon), HALF * (thiz->sigma + other->sigma), thiz->lambda);
        return 1;
      }
      return 0;
    default: return 0;
  }
}
/* modelClass_pair.f90:120-140 */
static opair_t pair_container_mix(const opair_t *a, const opair_t *b) {
  opair_t c; memset(&c, 0, sizeof c);
  if (!pair_mix_rule(&b->model, &a->model, &c.model) && !pair_mix_rule(&a->model, &b->model, &c.model)) {
    memset(&c.model, 0, sizeof c.model);
    c.model.kind = K_PAIR_NONE;
    char msg[256];
    snprintf(msg, sizeof msg, "no mixing rule found for models %s and %s", model_name(&a->model), model_name(&b->model));
    owarning(msg);
  }
  c.coulomb = a->coulomb && b->coulomb;
  if (c.coulomb) c.kCoul = sqrt(a->kCoul * b->kCoul);
  return c;
}

/* ============================================================ system state */

#define ME(md) ((odata_t *)(md).Data)
#define PAIR(me, i, j, l) ((me)->pair[((i) - 1) + (me)->ntypes * (((j) - 1) + (me)->ntypes * ((l) - 1))])
#define T2(me, arr, i, j) ((me)->arr[((i) - 1) + (me)->ntypes * ((j) - 1)])

static void list_alloc(olist_t *l, int nitems, int nlists) {
  l->nitems = nitems; l->count = 0;
  l->first = xcalloc(nlists + 1, sizeof(int));
  l->middle = xcalloc(nlists + 1, sizeof(int));
  l->last = xcalloc(nlists + 1, sizeof(int));
  l->item = xcalloc(nitems + 1, sizeof(int));
  l->value = xcalloc(nitems + 1, sizeof(double));
}
static void list_resize(olist_t *l, int nitems) {
  l->item = realloc(l->item, (nitems + 1) * sizeof(int));
  l->value = realloc(l->value, (nitems + 1) * sizeof(double));
  if (!l->item || !l->value) oerror("oracle", "out of memory");
  l->nitems = nitems;
}

/* ---------------------------------------------------------- rigid bodies */

/* ArBee.f90:156-176: B(q), C(q) are 4x3; Bt, Ct are their transposes. */
static void matB(const double *q, double *B) {  /* B[r*3+c], r in 0..3 */
  double v[12] = { -q[1], -q[2], -q[3],
                    q[0], -q[3],  q[2],
                    q[3],  q[0], -q[1],
                   -q[2],  q[1],  q[0] };
  memcpy(B, v, sizeof v);
}
static void matC(const double *q, double *C) {
  double v[12] = { -q[1], -q[2], -q[3],
                    q[0],  q[3], -q[2],
                   -q[3],  q[0],  q[1],
                    q[2], -q[1],  q[0] };
  memcpy(C, v, sizeof v);
}
static void mul43(const double *M, const double *x, double *y) {   /* y(4) = M(4x3) x(3) */
  for (int r = 0; r < 4; r++) y[r] = M[3 * r] * x[0] + M[3 * r + 1] * x[1] + M[3 * r + 2] * x[2];
}
static void mul34t(const double *M, const double *x, double *y) {  /* y(3) = M^T x(4) */
  for (int c = 0; c < 3; c++) y[c] = M[c] * x[0] + M[3 + c] * x[1] + M[6 + c] * x[2] + M[9 + c] * x[3];
}

/* ArBee.f90:77-93 */
static void body_setup(obody_t *b, const int *indexes, int n, const double *masses) {
  b->NP = n; b->dof = 6;
  b->index = xcalloc(n, sizeof(int));
  b->M = xcalloc(n, sizeof(double));
  b->d = xcalloc(3 * n, sizeof(double));
  b->delta = xcalloc(3 * n, sizeof(double));
  double s = 0;
  for (int k = 0; k < n; k++) { b->index[k] = indexes[k]; b->M[k] = masses[k]; s += b->M[k]; }
  b->mass = s;
  b->invMass = ONE / b->mass;
}

/* ArBee.f90:97-134 */
static void body_update(obody_t *b, const double *coords) {
  int n = b->NP;
  double inertia[9] = {0}, A[9], At[9], w[3];
  for (int x = 0; x < 3; x++) {
    double s = 0;
    for (int k = 0; k < n; k++) s += b->M[k] * coords[3 * k + x];
    b->rcm[x] = s * b->invMass;
  }
  for (int k = 0; k < n; k++) for (int x = 0; x < 3; x++) b->delta[3 * k + x] = coords[3 * k + x] - b->rcm[x];
  double s11 = 0, s22 = 0, s33 = 0, s12 = 0, s13 = 0, s23 = 0;
  for (int k = 0; k < n; k++) {
    const double *d = b->delta + 3 * k;
    s11 += b->M[k] * (d[1] * d[1] + d[2] * d[2]);
    s22 += b->M[k] * (d[0] * d[0] + d[2] * d[2]);
    s33 += b->M[k] * (d[0] * d[0] + d[1] * d[1]);
    s12 += b->M[k] * d[0] * d[1];
    s13 += b->M[k] * d[0] * d[2];
    s23 += b->M[k] * d[1] * d[2];
  }
  AM(inertia,1,1) = s11; AM(inertia,2,2) = s22; AM(inertia,3,3) = s33;
  AM(inertia,1,2) = -s12; AM(inertia,1,3) = -s13; AM(inertia,2,3) = -s23;
  diagonalization(inertia, A, w);
  for (int r = 1; r <= 3; r++) for (int c = 1; c <= 3; c++) AM(At, r, c) = AM(A, c, r);
  memcpy(A, At, sizeof A);
  for (int k = 0; k < 3; k++) { b->MoI[k] = w[k]; b->invMoI[k] = ONE / b->MoI[k]; }
  double I1 = b->MoI[0], I2 = b->MoI[1], I3 = b->MoI[2];
  b->I113 = ONE / (I1 * (I1 - I3));
  b->I313 = ONE / (I3 * (I1 - I3));
  b->I223 = ONE / (I2 * (I2 - I3));
  b->I212 = ONE / (I2 * (I1 - I2));
  b->m312 = (I3 - I1) / I2;
  quaternion_of(A, b->q);
  for (int k = 0; k < n; k++)
    for (int r = 1; r <= 3; r++)
      b->d[3 * k + r - 1] = AM(A, r, 1) * b->delta[3 * k] + AM(A, r, 2) * b->delta[3 * k + 1] + AM(A, r, 3) * b->delta[3 * k + 2];
}

/* delta = Ct(q) * (B(q) * d): ArBee.f90:193, :320 */
static void body_delta_from_q(obody_t *b) {
  double B[12], C[12];
  matB(b->q, B); matC(b->q, C);
  for (int k = 0; k < b->NP; k++) {
    double t[4];
    mul43(B, b->d + 3 * k, t);
    mul34t(C, t, b->delta + 3 * k);
  }
}

/* ArBee.f90:200-221 */
static void rotate_uniaxial(obody_t *b, int k, double dt) {
  double BkQ[4], BkPi[4], om, vsin, vcos;
  const double *q = b->q, *p = b->pi;
  switch (k) {
    case 1:
      BkQ[0] = -q[1]; BkQ[1] = q[0]; BkQ[2] = q[3]; BkQ[3] = -q[2];
      BkPi[0] = -p[1]; BkPi[1] = p[0]; BkPi[2] = p[3]; BkPi[3] = -p[2];
      break;
    case 2:
      BkQ[0] = -q[2]; BkQ[1] = -q[3]; BkQ[2] = q[0]; BkQ[3] = q[1];
      BkPi[0] = -p[2]; BkPi[1] = -p[3]; BkPi[2] = p[0]; BkPi[3] = p[1];
      break;
    default:
      BkQ[0] = -q[3]; BkQ[1] = q[2]; BkQ[2] = -q[1]; BkQ[3] = q[0];
      BkPi[0] = -p[3]; BkPi[1] = p[2]; BkPi[2] = -p[1]; BkPi[3] = p[0];
  }
  om = dt * (p[0] * BkQ[0] + p[1] * BkQ[1] + p[2] * BkQ[2] + p[3] * BkQ[3]) / (4.0 * b->MoI[k - 1]);
  vsin = sin(om); vcos = cos(om);
  for (int i = 0; i < 4; i++) { b->q[i] = vcos * b->q[i] + vsin * BkQ[i]; b->pi[i] = vcos * b->pi[i] + vsin * BkPi[i]; }
}

/* ArBee.f90:180-196 */
static void rotate_no_squish(obody_t *b, double delta_t, int n) {
  double dt = delta_t / n, half_dt = HALF * dt;
  for (int i = 1; i <= n; i++) {
    rotate_uniaxial(b, 3, half_dt);
    rotate_uniaxial(b, 2, half_dt);
    rotate_uniaxial(b, 1, dt);
    rotate_uniaxial(b, 2, half_dt);
    rotate_uniaxial(b, 3, half_dt);
  }
  body_delta_from_q(b);
}

/* ArBee.f90:292-297 */
static double Theta(double x, double n, double m) {
  double x2 = x * x;
  return -THIRD * n * x * x2 * Carlson_RJ(ONE - x2, ONE - m * x2, ONE, ONE + n * x2);
}
/* ArBee.f90:299-310 */
static double deltaFcn(double u0, double c0, double s0, double d0, double u, double cn, double sn, double dn,
                       double m, int jump, double alpha) {
  double eta = alpha * alpha / (ONE - alpha * alpha), C = sqrt(m + eta);
  double r = u - u0 + sign1(ONE, cn) * Theta(sn, eta, m) - sign1(ONE, c0) * Theta(s0, eta, m)
             + (alpha / C) * (atan(C * sn / dn) - atan(C * s0 / d0));
  if (jump != 0) r = r + jump * TWO * Theta(ONE, eta, m);
  return (eta + ONE) * r;
}
/* ArBee.f90:312-324 */
static double deltaFdn(double u0, double c0, double s0, double u, double cn, double sn,
                       double m, int jump, double alpha) {
  double eta = alpha * alpha;
  eta = eta / (ONE - eta);
  double k2eta = m * eta, C = sqrt(ONE + k2eta);
  double r = u - u0 + sign1(ONE, cn) * Theta(sn, k2eta, m) - sign1(ONE, c0) * Theta(s0, k2eta, m)
             + (alpha / C) * (atan(C * sn / cn) - atan(C * s0 / c0));
  if (jump != 0) r = r + jump * (TWO * Theta(ONE, k2eta, m) + (alpha / C) * PI_);
  return (eta + ONE) * r;
}

/* Exact torque-free rotation of an asymmetric top: ArBee.f90:225-290 */
static void rotate_exact(obody_t *b, double dt) {
  int i0, jump;
  double w0[3], Iw[3], Lsq, TwoKr, r1, r3, l1, l3, lmin, a[3], m, K, wp, s0, c0, u0, u;
  double L, deltaF, phi, z0[4], z[4], jac[3], inv2K;
  for (int k = 0; k < 3; k++) { w0[k] = b->omega[k]; Iw[k] = b->MoI[k] * w0[k]; }
  Lsq = Iw[1] * Iw[1] + Iw[2] * Iw[2];
  if (Lsq < EPS_ONE) { rotate_uniaxial(b, 1, dt); return; }
  Lsq = Iw[0] * Iw[0] + Lsq;
  L = sqrt(Lsq);
  TwoKr = Iw[0] * w0[0] + Iw[1] * w0[1] + Iw[2] * w0[2];
  r1 = Lsq - TwoKr * b->MoI[2];
  r3 = TwoKr * b->MoI[0] - Lsq;
  l1 = b->I223 * r1;
  l3 = b->I212 * r3;
  lmin = fmin(l1, l3);
  a[0] = sign1(ONE, w0[0]) * sqrt(b->I113 * r1);
  a[1] = sqrt(lmin);
  a[2] = sign1(ONE, w0[2]) * sqrt(b->I313 * r3);
  m = lmin / fmax(l1, l3);
  K = Carlson_RF(ZERO, ONE - m, ONE);
  inv2K = HALF / K;
  s0 = w0[1] / a[1];
  if (fabs(s0) < ONE) {
    c0 = (l1 < l3) ? w0[0] / a[0] : w0[2] / a[2];
    u0 = s0 * Carlson_RF(ONE - s0 * s0, ONE - m * s0 * s0, ONE);
    i0 = staircase(u0 * inv2K);
  } else {
    a[1] = fabs(w0[1]);
    s0 = sign1(ONE, s0);
    c0 = ZERO;
    u0 = sign1(K, s0);
    i0 = 0;
  }
  wp = b->m312 * a[0] * a[2] / a[1];
  u = wp * dt + u0;
  jump = staircase(u * inv2K) - i0;
  jacobi(u, m, jac);
  double sn = jac[0], cn = jac[1], dn = jac[2];
  if (l1 < l3) {
    b->omega[0] = a[0] * cn; b->omega[1] = a[1] * sn; b->omega[2] = a[2] * dn;
    deltaF = deltaFcn(u0, c0, s0, w0[2] / a[2], u, cn, sn, dn, m, jump, b->MoI[0] * a[0] / L);
  } else {
    b->omega[0] = a[0] * dn; b->omega[1] = a[1] * sn; b->omega[2] = a[2] * cn;
    deltaF = deltaFdn(u0, c0, s0, u, cn, sn, m, jump, b->MoI[0] * a[0] / L);
  }
  phi = (Lsq * (u - u0) + r3 * deltaF) / (TWO * L * b->MoI[0] * wp);
  z0[0] = Iw[2]; z0[1] = Iw[1]; z0[2] = L - Iw[0]; z0[3] = ZERO;
  for (int k = 0; k < 3; k++) Iw[k] = b->MoI[k] * b->omega[k];
  double cph = cos(phi), sph = sin(phi);
  z[0] = Iw[2] * cph - Iw[1] * sph;
  z[1] = Iw[1] * cph + Iw[2] * sph;
  z[2] = (L - Iw[0]) * cph + ZERO * sph;
  z[3] = ZERO * cph + (L - Iw[0]) * sph;
  /* q = normalize( z*(z0.q) + C(z) (Ct(z0) q) ) */
  double Cz[12], Cz0[12], t3[3], t4[4], qn[4];
  matC(z, Cz); matC(z0, Cz0);
  mul34t(Cz0, b->q, t3);
  mul43(Cz, t3, t4);
  double dq = z0[0] * b->q[0] + z0[1] * b->q[1] + z0[2] * b->q[2] + z0[3] * b->q[3];
  for (int k = 0; k < 4; k++) qn[k] = z[k] * dq + t4[k];
  double nrm = sqrt(qn[0] * qn[0] + qn[1] * qn[1] + qn[2] * qn[2] + qn[3] * qn[3]);
  for (int k = 0; k < 4; k++) b->q[k] = qn[k] / nrm;
  double B[12], twoIw[3] = { TWO * Iw[0], TWO * Iw[1], TWO * Iw[2] };
  matB(b->q, B);
  mul43(B, twoIw, b->pi);
  body_delta_from_q(b);
}

/* ArBee.f90:328-336 */
static void body_particle_momenta(const obody_t *b, double *P /* 3*NP */) {
  double B[12], C[12], At[9], omega[3];
  matB(b->q, B); matC(b->q, C);
  /* At = Ct(q) * B(q): (3x4)(4x3) */
  for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) {
    double s = 0; for (int k = 0; k < 4; k++) s += C[3 * k + r] * B[3 * k + c];
    At[3 * r + c] = s;
  }
  for (int r = 0; r < 3; r++) omega[r] = At[3 * r] * b->omega[0] + At[3 * r + 1] * b->omega[1] + At[3 * r + 2] * b->omega[2];
  for (int k = 0; k < b->NP; k++) {
    double cr[3];
    cross3(omega, b->delta + 3 * k, cr);
    for (int x = 0; x < 3; x++) P[3 * k + x] = b->M[k] * (b->invMass * b->pcm[x] + cr[x]);
  }
}

/* ArBee.f90:340-352 */
static void body_force_and_torque(obody_t *b, const double *F) {
  for (int x = 0; x < 3; x++) { b->F[x] = ZERO; b->tau[x] = ZERO; }
  for (int j = 0; j < b->NP; j++) {
    const double *Fj = F + 3 * (b->index[j] - 1);
    double cr[3];
    for (int x = 0; x < 3; x++) b->F[x] = b->F[x] + Fj[x];
    cross3(b->delta + 3 * j, Fj, cr);
    for (int x = 0; x < 3; x++) b->tau[x] = b->tau[x] + cr[x];
  }
}

/* ArBee.f90:356-366 */
static void body_assign_momenta(obody_t *b, const double *input, int size) {
  if (size == 3) {
    double B[12], t[3];
    for (int k = 0; k < 3; k++) b->omega[k] = input[k];
    matB(b->q, B);
    for (int k = 0; k < 3; k++) t[k] = TWO * b->MoI[k] * b->omega[k];
    mul43(B, t, b->pi);
  } else {
    double B[12], t[3];
    for (int k = 0; k < 4; k++) b->pi[k] = input[k];
    matB(b->q, B);
    mul34t(B, b->pi, t);
    for (int k = 0; k < 3; k++) b->omega[k] = HALF * b->invMoI[k] * t[k];
  }
}

/* ------------------------------------------------------------ EmDee_system */

static double *layer_energy_ptr(odata_t *me, int layer) { return (double *)((char *)me->layerEnergy + (size_t)(layer - 1) * sizeof(oenergy_t)); }

/* EmDeeData.f90:268-355 */
static void allocate_rigid_bodies(odata_t *me, const int *bodies) {
  int N = me->natoms;
  me->atomBody = xcalloc(N + 1, sizeof(int));
  if (bodies) {
    /* clean_body_indices: EmDeeData.f90:312-352 */
    int *index = xcalloc(N + 1, sizeof(int)), *saved = xcalloc(N + 1, sizeof(int));
    int *amount = xcalloc(N + 1, sizeof(int)), *first = xcalloc(N + 1, sizeof(int));
    int n = 0;
    for (int i = 1; i <= N; i++) {
      int ibody = bodies[i - 1];
      if (ibody > 0) {
        int j = 0, not_found = 1;
        while (not_found && j < n) { j = j + 1; not_found = saved[j] != ibody; }
        if (not_found) { n = n + 1; amount[n] = 1; saved[n] = ibody; first[n] = i; index[i] = 0; }
        else { amount[j] = amount[j] + 1; index[i] = j; index[first[j]] = j; }
      } else index[i] = 0;
    }
    int i = 0;
    for (int j = 1; j <= n; j++) if (amount[j] > 1) { i = i + 1; saved[j] = i; }
    for (int k = 1; k <= N; k++) if (index[k] > 0) index[k] = saved[index[k]];
    int nb = 0;
    for (int k = 1; k <= N; k++) { me->atomBody[k] = index[k]; if (index[k] > nb) nb = index[k]; }
    me->nbodies = nb;
    me->free_ = xcalloc(N + 1, sizeof(int));
    me->nfree = 0;
    for (int k = 1; k <= N; k++) if (me->atomBody[k] == 0) me->free_[++me->nfree] = k;
    me->body = xcalloc(nb + 1, sizeof(obody_t));
    int *ind = xcalloc(N + 1, sizeof(int));
    double *ms = xcalloc(N + 1, sizeof(double));
    for (int b = 1; b <= nb; b++) {
      int c = 0;
      for (int k = 1; k <= N; k++) if (me->atomBody[k] == b) { ind[c] = k; ms[c] = me->mass[k]; c++; }
      body_setup(&me->body[b], ind, c, ms);
    }
    free(ind); free(ms);
    i = me->nbodies;
    for (int j = 1; j <= N; j++) if (me->atomBody[j] == 0) { i = i + 1; me->atomBody[j] = i; }
    free(index); free(saved); free(amount); free(first);
  } else {
    me->nbodies = 0;
    me->free_ = xcalloc(N + 1, sizeof(int));
    for (int k = 1; k <= N; k++) { me->free_[k] = k; me->atomBody[k] = k; }
    me->nfree = N;
    me->body = xcalloc(1, sizeof(obody_t));
  }
  me->threadFreeAtoms = (me->nfree + me->nthreads - 1) / me->nthreads;
  me->threadBodies = (me->nbodies + me->nthreads - 1) / me->nthreads;
}

/* EmDeeCode.f90:69-208 */
tEmDee EmDee_system(int threads, int layers, double rc, double skin, int N, int *types, double *masses, int *bodies) {
  tEmDee md; memset(&md, 0, sizeof md);
  odata_t *me = xcalloc(1, sizeof *me);
  if (threads < 1) threads = 1;
  me->nthreads = threads;
  me->nlayers = layers;
  me->Rc = rc; me->skin = skin; me->RcSq = rc * rc;
  me->xRc = rc + skin; me->xRcSq = me->xRc * me->xRc;
  me->InRc = me->Rc; me->InRcSq = me->RcSq; me->xInRcSq = me->xRcSq;
  me->skinSq = skin * skin;
  me->natoms = N;
  me->threadAtoms = (N + threads - 1) / threads;
  me->kspace_active = 0;
  me->atomType = xcalloc(N + 1, sizeof(int));
  if (types) {
    int mn = types[0], mx = types[0];
    for (int i = 0; i < N; i++) { if (types[i] < mn) mn = types[i]; if (types[i] > mx) mx = types[i]; }
    if (mn != 1) oerror("system setup", "wrong specification of atom types");
    me->ntypes = mx;
    for (int i = 1; i <= N; i++) me->atomType[i] = types[i - 1];
  } else {
    me->ntypes = 1;
    for (int i = 1; i <= N; i++) me->atomType[i] = 1;
  }
  me->mass = xcalloc(N + 1, sizeof(double));
  me->invMass = xcalloc(N + 1, sizeof(double));
  if (masses) {
    /* masses are per type; the reference indexes them with ptype(:) */
    double tm = 0;
    for (int i = 1; i <= N; i++) {
      double m = masses[me->atomType[i] - 1];
      me->mass[i] = m; me->invMass[i] = ONE / m; tm += m;
    }
    me->totalMass = tm;
  } else {
    for (int i = 1; i <= N; i++) { me->mass[i] = ONE; me->invMass[i] = ONE; }
    me->totalMass = (double)N;
  }
  me->startTime = wtime();
  me->P = xcalloc(3 * N, sizeof(double));
  me->R0 = xcalloc(3 * N, sizeof(double));
  me->charge = xcalloc(N + 1, sizeof(double));
  me->charged = xcalloc(N + 1, sizeof(int));
  me->atomCell = xcalloc(N + 1, sizeof(int));
  me->Rowned = me->Powned = me->bodyOwned = me->Lowned = 1;
  allocate_rigid_bodies(me, bodies);
  me->cellItem = xcalloc(N + 1, sizeof(int));
  me->neighbor = xcalloc(threads, sizeof(olist_t));
  for (int t = 0; t < threads; t++) list_alloc(&me->neighbor[t], 2000, N);
  me->exCount = xcalloc(N + 1, sizeof(int));
  me->exCap = xcalloc(N + 1, sizeof(int));
  me->exItem = xcalloc(N + 1, sizeof(int *));
  int nt = me->ntypes;
  me->pair = xcalloc((size_t)nt * nt * layers, sizeof(opair_t));
  for (size_t k = 0; k < (size_t)nt * nt * layers; k++) me->pair[k].model.kind = K_PAIR_NONE;
  me->multilayer = xcalloc(nt * nt, sizeof(int));
  me->overridable = xcalloc(nt * nt, sizeof(int));
  for (int k = 0; k < nt * nt; k++) me->overridable[k] = 1;
  me->interact = xcalloc(nt * nt, sizeof(int));
  me->coul = xcalloc(layers, sizeof(omodel_t));
  for (int l = 0; l < layers; l++) me->coul[l].kind = K_COUL_NONE;
  me->multilayer_coulomb = 0;
  me->pairs_exist = xcalloc(layers + 1, sizeof(int));
  me->useInRc = xcalloc(layers + 1, sizeof(int));
  me->bonded = xcalloc(layers + 1, sizeof(int));
  for (int l = 1; l <= layers; l++) me->bonded[l] = 1;
  me->forcesUpToDate = xcalloc(layers + 1, sizeof(int));
  me->layerF = xcalloc((size_t)3 * N * layers, sizeof(double));
  me->layer = 1;
  me->F = me->layerF;
  me->initialized = 0;
  me->random.seeding_required = 1;
  me->layerEnergy = xcalloc(layers, sizeof(oenergy_t));
  me->layerVirial = xcalloc(layers, sizeof(ovirial_t));
  md.Builds = 0;
  md.DoF = 3 * (N - 1);
  md.RotDoF = 0;
  md.Data = me;
  md.Options.Translate = true;
  md.Options.Rotate = true;
  md.Options.RotationMode = 0;
  md.Options.AutoBodyUpdate = true;
  md.Options.Compute = true;
  return md;
}

/* EmDeeCode.f90:212-235 */
void *EmDee_memory_address(tEmDee md, char *option) {
  odata_t *me = ME(md);
  if (!strcmp(option, "coordinates")) return me->R;
  if (!strcmp(option, "momenta")) return me->P;
  if (!strcmp(option, "forces")) return me->F;
  if (!strcmp(option, "layerForces")) return me->layerF;
  char msg[300]; snprintf(msg, sizeof msg, "invalid option %s", option);
  oerror("memory address retrieving", msg);
  return NULL;
}

/* EmDeeCode.f90:239-269 */
void EmDee_share_phase_space(tEmDee mdkeep, tEmDee *mdlose) {
  const char *task = "phase space sharing";
  odata_t *keep = ME(mdkeep), *lose = ME(*mdlose);
  if (!(keep->initialized && lose->initialized)) oerror(task, "EmDee system 1 has not been initialized");
  if (keep->natoms != lose->natoms) oerror(task, "different numbers of atoms");
  for (int i = 1; i <= keep->natoms; i++) {
    if (keep->atomType[i] != lose->atomType[i]) oerror(task, "atom types do not match");
    if (keep->mass[i] != lose->mass[i]) oerror(task, "atom masses do not match");
    if (keep->atomBody[i] != lose->atomBody[i]) oerror(task, "rigid bodies do not match");
  }
  lose->R = keep->R; lose->P = keep->P; lose->body = keep->body; lose->Lbox = keep->Lbox;
  lose->Rowned = lose->Powned = lose->bodyOwned = lose->Lowned = 0;
  mdlose->Kinetic.Kinetic = mdkeep.Kinetic.Kinetic;
  for (int k = 0; k < 3; k++) { mdlose->Kinetic.TransPart[k] = mdkeep.Kinetic.TransPart[k]; mdlose->Kinetic.RotPart[k] = mdkeep.Kinetic.RotPart[k]; }
  mdlose->Kinetic.Rotational = mdkeep.Kinetic.Rotational;
}

/* EmDeeCode.f90:273-305 */
void EmDee_layer_based_parameters(tEmDee md, double InternalRc, int *Apply, int *Bonded) {
  const char *task = "layer-based parameter setting";
  odata_t *me = ME(md);
  if (me->initialized) oerror(task, "system has already been initialized");
  int any = 0;
  for (int l = 1; l <= me->nlayers; l++) { me->useInRc[l] = Apply[l - 1] != 0; any |= me->useInRc[l]; }
  if (any && (InternalRc <= ZERO || InternalRc > me->Rc)) oerror(task, "invalid internal cutoff specification");
  me->InRc = InternalRc;
  me->InRcSq = InternalRc * InternalRc;
  me->xInRcSq = (InternalRc + me->skin) * (InternalRc + me->skin);
  for (int l = 1; l <= me->nlayers; l++) me->bonded[l] = Bonded[l - 1] != 0;
  for (int l = 1; l <= me->nlayers; l++) coul_cutoff_setup(&me->coul[l - 1], me->useInRc[l] ? me->InRc : me->Rc);
}

/* EmDeeData.f90:227-264 */
static void set_pair_type(odata_t *me, int itype, int jtype, int layer, const omodel_t *container, double kCoul) {
  double layerRc = me->useInRc[layer] ? me->InRc : me->Rc;
  if (itype == jtype) {
    opair_t *p = &PAIR(me, itype, itype, layer);
    p->model = *container;
    p->coulomb = kCoul != ZERO;
    if (p->coulomb) p->kCoul = kCoul;
    modifier_setup(&p->model, layerRc);
    for (int ktype = 1; ktype <= me->ntypes; ktype++) {
      if (ktype != itype && T2(me, overridable, itype, ktype)) {
        PAIR(me, itype, ktype, layer) = pair_container_mix(&PAIR(me, ktype, ktype, layer), &PAIR(me, itype, itype, layer));
        modifier_setup(&PAIR(me, itype, ktype, layer).model, layerRc);
        PAIR(me, ktype, itype, layer) = PAIR(me, itype, ktype, layer);
      }
    }
  } else {
    opair_t *p = &PAIR(me, itype, jtype, layer);
    p->model = *container;
    p->coulomb = kCoul != ZERO;
    if (p->coulomb) p->kCoul = kCoul;
    modifier_setup(&p->model, layerRc);
    PAIR(me, jtype, itype, layer) = *p;
  }
}

/* EmDeeCode.f90:309-359 */
void EmDee_set_pair_model(tEmDee md, int itype, int jtype, void *model, double kCoul) {
  const char *task = "pair model setup";
  odata_t *me = ME(md);
  if (me->initialized) oerror(task, "cannot set model after coordinates have been defined");
  int ij[2] = { itype, jtype };
  if (!ranged(ij, 2, me->ntypes)) oerror(task, "provided type index is out of range");
  if (!model) oerror(task, "a valid pair model must be provided");
  omodel_t *c = model;
  if (!is_pair(c)) oerror(task, "a valid pair model must be provided");
  for (int l = 1; l <= me->nlayers; l++) set_pair_type(me, itype, jtype, l, c, kCoul);
  T2(me, multilayer, itype, jtype) = 0;
  T2(me, multilayer, jtype, itype) = 0;
  if (itype == jtype) {
    for (int k = 1; k <= me->ntypes; k++)
      if (k != itype && T2(me, overridable, itype, k)) {
        T2(me, multilayer, itype, k) = T2(me, multilayer, k, k);
        T2(me, multilayer, k, itype) = T2(me, multilayer, k, k);
      }
  } else {
    T2(me, overridable, itype, jtype) = 0;
    T2(me, overridable, jtype, itype) = 0;
  }
}

/* EmDeeCode.f90:363-412 */
void EmDee_set_pair_multimodel(tEmDee md, int itype, int jtype, void *model[], double kCoul[]) {
  const char *task = "pair multimodel setup";
  odata_t *me = ME(md);
  if (me->initialized) oerror(task, "cannot set model after coordinates have been defined");
  int ij[2] = { itype, jtype };
  if (!ranged(ij, 2, me->ntypes)) oerror(task, "provided type index is out of range");
  for (int l = 1; l <= me->nlayers; l++) {
    if (!model[l - 1]) { char msg[64]; snprintf(msg, sizeof msg, "%d valid pair models must be provided", me->nlayers); oerror(task, msg); }
    omodel_t *c = model[l - 1];
    if (!is_pair(c)) oerror(task, "a valid pair model must be provided");
    set_pair_type(me, itype, jtype, l, c, kCoul[l - 1]);
  }
  T2(me, multilayer, itype, jtype) = 1;
  T2(me, multilayer, jtype, itype) = 1;
  if (itype == jtype) {
    for (int k = 1; k <= me->ntypes; k++)
      if (k != itype && T2(me, overridable, itype, k)) { T2(me, multilayer, itype, k) = 1; T2(me, multilayer, k, itype) = 1; }
  } else {
    T2(me, overridable, itype, jtype) = 0;
    T2(me, overridable, jtype, itype) = 0;
  }
}

/* EmDeeCode.f90:416-444 */
void EmDee_set_kspace_model(tEmDee md, void *model) {
  const char *task = "kspace model setup";
  odata_t *me = ME(md);
  if (me->initialized) oerror(task, "cannot set model after coordinates have been defined");
  if (!model) oerror(task, "a valid kspace model must be provided");
  omodel_t *c = model;
  if (c->kind != K_KSPACE_EWALD) oerror(task, "a valid kspace model must be provided");
  me->kspace = xcalloc(1, sizeof(okspace_t));
  me->kspace->base = *c;
  me->kspace->nmax[0] = me->kspace->nmax[1] = me->kspace->nmax[2] = -1;
  me->kspace_active = 1;
}

/* EmDeeCode.f90:448-482 */
void EmDee_set_coul_model(tEmDee md, void *model) {
  const char *task = "coulomb model setup";
  odata_t *me = ME(md);
  if (me->initialized) oerror(task, "cannot set model after coordinates have been defined");
  if (!model) oerror(task, "a valid coulomb model must be provided");
  omodel_t *c = model;
  if (!is_coul(c)) oerror(task, "a valid coulomb model must be provided");
  for (int l = 1; l <= me->nlayers; l++) {
    double layerRc = me->useInRc[l] ? me->InRc : me->Rc;
    me->coul[l - 1] = *c;
    coul_cutoff_setup(&me->coul[l - 1], layerRc);
    modifier_setup(&me->coul[l - 1], layerRc);
  }
  me->multilayer_coulomb = 0;
}

/* EmDeeCode.f90:486-520 */
void EmDee_set_coul_multimodel(tEmDee md, void *model[]) {
  const char *task = "coulomb multimodel setup";
  odata_t *me = ME(md);
  if (me->initialized) oerror(task, "cannot set model after coordinates have been defined");
  for (int l = 1; l <= me->nlayers; l++) {
    double layerRc = me->useInRc[l] ? me->InRc : me->Rc;
    if (!model[l - 1]) { char msg[64]; snprintf(msg, sizeof msg, "%d valid coulomb models must be provided", me->nlayers); oerror(task, msg); }
    omodel_t *c = model[l - 1];
    if (!is_coul(c)) oerror(task, "a valid coulomb model must be provided");
    me->coul[l - 1] = *c;
    coul_cutoff_setup(&me->coul[l - 1], layerRc);
    modifier_setup(&me->coul[l - 1], layerRc);
  }
  me->multilayer_coulomb = 1;
}

/* EmDeeCode.f90:524-570: sorted, duplicate-free exclusion list per atom */
static void ex_add(odata_t *me, int i, int j) {
  int n = me->exCount[i];
  for (int k = 0; k < n; k++) if (me->exItem[i][k] == j) return;
  if (n == me->exCap[i]) {
    me->exCap[i] = me->exCap[i] ? 2 * me->exCap[i] : 4;
    me->exItem[i] = realloc(me->exItem[i], me->exCap[i] * sizeof(int));
  }
  int k = n;
  while (k > 0 && me->exItem[i][k - 1] > j) { me->exItem[i][k] = me->exItem[i][k - 1]; k--; }
  me->exItem[i][k] = j;
  me->exCount[i] = n + 1;
}
void EmDee_ignore_pair(tEmDee md, int i, int j) {
  odata_t *me = ME(md);
  int ij[2] = { i, j };
  if (i != j && ranged(ij, 2, me->natoms)) { ex_add(me, i, j); ex_add(me, j, i); }
}

static void struct_add(ostruct_t **arr, int *n, int *cap, int i, int j, int k, int l, const omodel_t *m) {
  if (*n == *cap) { *cap = *cap ? 2 * *cap : 16; *arr = realloc(*arr, *cap * sizeof(ostruct_t)); }
  ostruct_t *s = &(*arr)[(*n)++];
  s->i = i; s->j = j; s->k = k; s->l = l; s->model = *m;
}
/* EmDeeCode.f90:574-597 */
void EmDee_add_bond(tEmDee md, int i, int j, void *model) {
  odata_t *me = ME(md);
  int ij[2] = { i, j };
  if (!ranged(ij, 2, me->natoms)) oerror("add_bond", "atom index out of range");
  if (!model) oerror("add_bond", "a valid model must be provided");
  omodel_t *m = model;
  if (m->kind != K_BOND_HARMONIC && m->kind != K_BOND_NONE) oerror("add_bond", "the provided model must be a bond model");
  struct_add(&me->bonds, &me->nbonds, &me->capbonds, i, j, 0, 0, m);
  EmDee_ignore_pair(md, i, j);
}
/* EmDeeCode.f90:600-626 */
void EmDee_add_angle(tEmDee md, int i, int j, int k, void *model) {
  odata_t *me = ME(md);
  int ijk[3] = { i, j, k };
  if (!ranged(ijk, 3, me->natoms)) oerror("add_angle", "atom index out of range");
  if (!model) oerror("add_angle", "a valid model must be provided");
  omodel_t *m = model;
  if (m->kind != K_ANGLE_HARMONIC && m->kind != K_ANGLE_NONE) oerror("add_angle", "the provided model must be an angle model");
  struct_add(&me->angles, &me->nangles, &me->capangles, i, j, k, 0, m);
  EmDee_ignore_pair(md, i, j);
  EmDee_ignore_pair(md, i, k);
  EmDee_ignore_pair(md, j, k);
}
/* EmDeeCode.f90:628-655 */
void EmDee_add_dihedral(tEmDee md, int i, int j, int k, int l, void *model) {
  odata_t *me = ME(md);
  int ijkl[4] = { i, j, k, l };
  if (!ranged(ijkl, 4, me->natoms)) oerror("add_dihedral", "atom index out of range");
  if (!model) oerror("add_dihedral", "a valid model must be provided");
  omodel_t *m = model;
  if (m->kind != K_DIHEDRAL_NONE && m->kind != K_DIHEDRAL_HARMONIC)
    oerror("add_dihedral", "the provided model must be a dihedral model");
  struct_add(&me->dihedrals, &me->ndihedrals, &me->capdihedrals, i, j, k, l, m);
  EmDee_ignore_pair(md, i, j); EmDee_ignore_pair(md, i, k); EmDee_ignore_pair(md, i, l);
  EmDee_ignore_pair(md, j, k); EmDee_ignore_pair(md, j, l); EmDee_ignore_pair(md, k, l);
}

/* ========================================================= neighbor lists */

/* The 62-cell half stencil of neighbor_lists.f90:26-35: every offset d in
   [-2,2]^3 in the upper half-space (dz > 0, or dz = 0 and dy > 0, or
   dz = dy = 0 and dx > 0), ordered by |d|^2, then dx, dy, dz. */
static int stencil[62][3];
static int stencil_ready = 0;
static void make_stencil(void) {
  if (stencil_ready) return;
  int n = 0;
  for (int r2 = 1; r2 <= 12; r2++)
    for (int dx = -2; dx <= 2; dx++)
      for (int dy = -2; dy <= 2; dy++)
        for (int dz = -2; dz <= 2; dz++) {
          if (dx * dx + dy * dy + dz * dz != r2) continue;
          int upper = dz > 0 || (dz == 0 && (dy > 0 || (dy == 0 && dx > 0)));
          if (!upper) continue;
          stencil[n][0] = dx; stencil[n][1] = dy; stencil[n][2] = dz; n++;
        }
  stencil_ready = 1;
}
void oracle_stencil(int *out) { make_stencil(); memcpy(out, stencil, sizeof stencil); }

/* neighbor_lists.f90:41-59 */
static double maximum_approach_sq(int N, const double *R, const double *R0) {
  double maximum, next, deltaSq, d[3];
  for (int x = 0; x < 3; x++) d[x] = R[x] - R0[x];
  maximum = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  next = maximum;
  for (int i = 2; i <= N; i++) {
    for (int x = 0; x < 3; x++) d[x] = R[3 * (i - 1) + x] - R0[3 * (i - 1) + x];
    deltaSq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (deltaSq > maximum) { next = maximum; maximum = deltaSq; }
  }
  return maximum + 2 * sqrt(maximum * next) + next;
}
double oracle_maximum_approach_sq(int N, const double *R, const double *R0) { return maximum_approach_sq(N, R, R0); }

static inline int cpbc(int x, int M) { return x < 0 ? x + M : (x >= M ? x - M : x); }

/* neighbor_lists.f90:63-171 */
static void distribute_atoms(odata_t *me, int M, const double *Rs) {
  int MM = M * M, N = me->natoms, T = me->nthreads;
  int make_cells = M != me->mcells;
  int cells_per_thread = 0;
  make_stencil();
  if (make_cells) {
    me->mcells = M;
    me->ncells = M * MM;
    if (me->ncells > me->maxcells) {
      free(me->cellNeighbor); free(me->cellFirst); free(me->cellLast); free(me->atomsInCell);
      me->cellNeighbor = xcalloc((size_t)(me->ncells + 1) * 62, sizeof(int));
      me->cellFirst = xcalloc(me->ncells + 1, sizeof(int));
      me->cellLast = xcalloc(me->ncells + 1, sizeof(int));
      me->atomsInCell = xcalloc(me->ncells + 1, sizeof(int));
      free(me->threadCellFirst); free(me->threadCellLast);
      me->threadCellFirst = xcalloc(T + 1, sizeof(int));
      me->threadCellLast = xcalloc(T + 1, sizeof(int));
      me->maxcells = me->ncells;
    }
    cells_per_thread = (me->ncells + T - 1) / T;
  }
  int *next = xcalloc(N + 1, sizeof(int));
  int *threadNatoms = xcalloc(T + 1, sizeof(int));
  int *maxN = xcalloc(T + 1, sizeof(int));
#pragma omp parallel num_threads(T)
  {
    int thread = thread_num() + 1, first, last;
    if (make_cells) {
      first = (thread - 1) * cells_per_thread + 1;
      last = me->ncells < thread * cells_per_thread ? me->ncells : thread * cells_per_thread;
      for (int icell = first; icell <= last; icell++) {
        int k = icell - 1, iz = k / MM, j = k - iz * MM, iy = j / M, ix = j - iy * M;
        for (int s = 0; s < 62; s++)
          me->cellNeighbor[(size_t)icell * 62 + s] = 1 + cpbc(ix + stencil[s][0], M) + cpbc(iy + stencil[s][1], M) * M +
                                                     cpbc(iz + stencil[s][2], M) * MM;
      }
      me->threadCellFirst[thread] = first;
      me->threadCellLast[thread] = last;
    } else {
      first = me->threadCellFirst[thread];
      last = me->threadCellLast[thread];
    }
    int a1 = (thread - 1) * me->threadAtoms + 1, aN = N < thread * me->threadAtoms ? N : thread * me->threadAtoms;
    for (int i = a1; i <= aN; i++) {
      int ic[3];
      for (int x = 0; x < 3; x++) {
        double s = Rs[3 * (i - 1) + x];
        ic[x] = (int)(M * (s - floor(s)));
        /* s - floor(s) can round to 1.0 for tiny negative s; the cell is
           then the periodic image 0 (the reference would index out of range). */
        if (ic[x] >= M) ic[x] -= M;
      }
      me->atomCell[i] = 1 + ic[0] + M * ic[1] + MM * ic[2];
    }
#pragma omp barrier
    int *head = xcalloc(last - first + 2, sizeof(int));
    for (int c = first; c <= last; c++) me->atomsInCell[c] = 0;
    for (int i = 1; i <= N; i++) {
      int icell = me->atomCell[i];
      if (icell >= first && icell <= last) {
        next[i] = head[icell - first];
        head[icell - first] = i;
        me->atomsInCell[icell]++;
      }
    }
    int s = 0;
    for (int c = first; c <= last; c++) s += me->atomsInCell[c];
    threadNatoms[thread] = s;
#pragma omp barrier
    int mx = 0, k = 0;
    for (int t = 1; t < thread; t++) k += threadNatoms[t];
    for (int icell = first; icell <= last; icell++) {
      me->cellFirst[icell] = k + 1;
      int i = head[icell - first];
      while (i != 0) { k = k + 1; me->cellItem[k] = i; i = next[i]; }
      me->cellLast[icell] = k;
      if (me->atomsInCell[icell] > mx) mx = me->atomsInCell[icell];
    }
    maxN[thread] = mx;
    free(head);
  }
  me->maxatoms = 0;
  for (int t = 1; t <= T; t++) if (maxN[t] > me->maxatoms) me->maxatoms = maxN[t];
  me->maxpairs = (me->maxatoms * ((2 * 62 + 1) * me->maxatoms - 1)) / 2;
  free(next); free(threadNatoms); free(maxN);
}

/* neighbor_lists.f90:199-300 */
static void build_neighbor_lists(odata_t *me, int thread, const double *Rs) {
  double invL2 = ONE / (*me->Lbox * *me->Lbox);
  double xRc2 = me->xRcSq * invL2, xInRc2 = me->xInRcSq * invL2;
  int N = me->natoms;
  char *include = xcalloc(N + 1, 1);
  memset(include, 1, N + 1);
  int *atom = xcalloc((size_t)(62 + 1) * me->maxatoms + 1, sizeof(int));
  double *rsq = xcalloc((size_t)(62 + 1) * me->maxatoms + 1, sizeof(double));
  olist_t *nb = &me->neighbor[thread - 1];
  int npairs = 0;
  for (int icell = me->threadCellFirst[thread]; icell <= me->threadCellLast[thread]; icell++) {
    int nlocal = me->atomsInCell[icell], ntotal = nlocal;
    const int *neigh = me->cellNeighbor + (size_t)icell * 62;
    for (int m = 0; m < 62; m++) ntotal += me->atomsInCell[neigh[m]];
    int c = 0;
    for (int p = me->cellFirst[icell]; p <= me->cellLast[icell]; p++) atom[++c] = me->cellItem[p];
    for (int m = 0; m < 62; m++)
      for (int p = me->cellFirst[neigh[m]]; p <= me->cellLast[neigh[m]]; p++) atom[++c] = me->cellItem[p];
    if (nb->nitems < npairs + nlocal * ntotal) list_resize(nb, npairs + nlocal * ntotal + 2000);
    for (int k = 1; k <= nlocal; k++) {
      int i = atom[k];
      int first = npairs + 1;
      nb->first[i] = first;
      int ipairs = 0, middle = 0;
      int itype = me->atomType[i], ibody = me->atomBody[i];
      for (int e = 0; e < me->exCount[i]; e++) include[me->exItem[i][e]] = 0;
      const double *Ri = Rs + 3 * (i - 1);
      for (int m = k + 1; m <= ntotal; m++) {
        const double *Rj = Rs + 3 * (atom[m] - 1);
        double d0 = Ri[0] - Rj[0], d1 = Ri[1] - Rj[1], d2 = Ri[2] - Rj[2];
        d0 = d0 - anint(d0); d1 = d1 - anint(d1); d2 = d2 - anint(d2);
        rsq[m] = d0 * d0 + d1 * d1 + d2 * d2;
      }
      int *item = nb->item + first;       /* item(first:) -> item[0..] */
      double *value = nb->value + first;
      for (int m = k + 1; m <= ntotal; m++) {
        double r2 = rsq[m];
        if (r2 < xRc2) {
          int j = atom[m];
          if (include[j] && me->atomBody[j] != ibody && T2(me, interact, me->atomType[j], itype)) {
            /* insert_neighbor: neighbor_lists.f90:285-298 */
            int kk;
            for (kk = ipairs; kk >= 1; kk--) {
              if (value[kk - 1] < r2) break;
              value[kk] = value[kk - 1];
              item[kk] = item[kk - 1];
            }
            value[kk] = r2;
            item[kk] = j;
            ipairs++;
            if (r2 < xInRc2) middle++;
          }
        }
      }
      for (int e = 0; e < me->exCount[i]; e++) include[me->exItem[i][e]] = 1;
      nb->middle[i] = npairs + middle;
      npairs = npairs + ipairs;
      nb->last[i] = npairs;
    }
  }
  nb->count = npairs;
  free(include); free(atom); free(rsq);
}

/* neighbor_lists.f90:175-195 */
static void handle_neighbor_lists(odata_t *me, int *builds, double *time, const double *Rs) {
  *time -= wtime();
  if (maximum_approach_sq(me->natoms, me->R, me->R0) > me->skinSq) {
    int M = (int)floor(2 * *me->Lbox / me->xRc);
    distribute_atoms(me, M > 5 ? M : 5, Rs);
    memcpy(me->R0, me->R, 3 * sizeof(double) * me->natoms);
    *builds += 1;
#pragma omp parallel num_threads(me->nthreads)
    build_neighbor_lists(me, thread_num() + 1, Rs);
  }
  *time += wtime();
}

/* Half-list pair set after the latest build: writes (i, j) pairs (1-based,
   i < j) and returns the count; `max` bounds the output. */
int oracle_neighbor_pairs(tEmDee md, int max, int *out) {
  odata_t *me = ME(md);
  int n = 0;
  for (int t = 1; t <= me->nthreads; t++) {
    olist_t *nb = &me->neighbor[t - 1];
    for (int icell = me->threadCellFirst[t]; icell <= me->threadCellLast[t]; icell++)
      for (int p = me->cellFirst[icell]; p <= me->cellLast[icell]; p++) {
        int i = me->cellItem[p];
        for (int m = nb->first[i]; m <= nb->last[i]; m++) {
          int j = nb->item[m];
          if (n < max) { out[2 * n] = i < j ? i : j; out[2 * n + 1] = i < j ? j : i; }
          n++;
        }
      }
  }
  return n;
}
int oracle_cells_per_side(tEmDee md) { return ME(md)->mcells; }

/* =========================================================== force terms */

static void apply_modifier(const omodel_t *m, int compute, double invR, double invR2, double *Eij, double *Wij, int is_p) {
  double rFc, r2fac, u, u2, u3, G, WG;
  switch (m->modifier) {
    case MOD_SHIFTED:
      if (compute) *Eij += m->eshift;
      break;
    case MOD_SHIFTED_FORCE:
      rFc = m->fshift / invR;
      *Wij = *Wij - rFc;
      if (compute) *Eij = *Eij + m->eshift + rFc;
      break;
    case MOD_SMOOTHED: case MOD_SHIFTED_SMOOTHED:
      r2fac = m->factor / invR;
      if (compute && (!legacy_shifted_smoothing || r2fac > m->Rm2fac)) *Eij = *Eij + m->eshift;
      if (r2fac > m->Rm2fac) {
        if (!compute) { *Eij = is_p ? pair_energy(m, invR, invR2) : coul_energy(m, invR, invR2); *Eij = *Eij + m->eshift; }
        u = r2fac - m->Rm2fac; u2 = u * u; u3 = u * u2;
        G = 1.0 + u3 * (15.0 * u - 6.0 * u2 - 10.0);
        WG = -30.0 * u2 * (2.0 * u - u2 - 1.0) * r2fac;
        *Wij = *Wij * G + *Eij * WG;
        *Eij = *Eij * G;
      }
      break;
    case MOD_SQUARE_SMOOTHED: case MOD_SHIFTED_SQUARE_SMOOTHED:
      r2fac = m->factor / invR2;
      if (compute && (!legacy_shifted_smoothing || r2fac > m->Rm2fac)) *Eij = *Eij + m->eshift;
      if (r2fac > m->Rm2fac) {
        if (!compute) { *Eij = is_p ? pair_energy(m, invR, invR2) : coul_energy(m, invR, invR2); *Eij = *Eij + m->eshift; }
        u = r2fac - m->Rm2fac; u2 = u * u; u3 = u * u2;
        G = 1.0 + u3 * (15.0 * u - 6.0 * u2 - 10.0);
        WG = -60.0 * u2 * (2.0 * u - u2 - 1.0) * r2fac;
        *Wij = *Wij * G + *Eij * WG;
        *Eij = *Eij * G;
      }
      break;
    default: break;
  }
}

/* EmDeeData.f90:644-685 + compute.f90:20-100 + apply_modifier.f90 */
static void compute_pairs(odata_t *me, int thread, int compute, const double *Rs, double *F,
                          double *Epair, double *Ecoul, double *Wpair, double *Wcoul) {
  int N = me->natoms;
  memset(F, 0, 3 * sizeof(double) * N);
  *Wpair = ZERO; *Wcoul = ZERO; *Epair = ZERO;
  if (!me->pairs_exist[me->layer]) return;
  double L2 = *me->Lbox * *me->Lbox, invL2 = ONE / L2, Rc2;
  olist_t *nb = &me->neighbor[thread - 1];
  const int *upper;
  if (me->useInRc[me->layer]) { Rc2 = me->InRcSq * invL2; upper = nb->middle; }
  else { Rc2 = me->RcSq * invL2; upper = nb->last; }
  const omodel_t *cm = &me->coul[me->layer - 1];
  int firstAtom = me->cellFirst[me->threadCellFirst[thread]];
  int lastAtom = me->cellLast[me->threadCellLast[thread]];
  for (int k = firstAtom; k <= lastAtom; k++) {
    int i = me->cellItem[k], itype = me->atomType[i];
    double Qi = me->charge[i];
    int icharged = me->charged[i];
    const double *Ri = Rs + 3 * (i - 1);
    double Fi[3] = { 0, 0, 0 };
    for (int m = nb->first[i]; m <= upper[i]; m++) {
      int j = nb->item[m];
      const double *Rj = Rs + 3 * (j - 1);
      double Rij[3];
      for (int x = 0; x < 3; x++) { double d = Ri[x] - Rj[x]; Rij[x] = d - anint(d); }
      double r2 = Rij[0] * Rij[0] + Rij[1] * Rij[1] + Rij[2] * Rij[2];
      if (r2 < Rc2) {
        double invR2 = invL2 / r2, invR = sqrt(invR2);
        int jtype = me->atomType[j];
        int ijcharged = icharged && me->charged[j];
        const opair_t *pair = &PAIR(me, jtype, itype, me->layer);
        double Eij = 0, Wij = 0, Wsum;
        if (compute) pair_compute(&pair->model, &Eij, &Wij, invR, invR2);
        else { double e; pair_compute(&pair->model, &e, &Wij, invR, invR2); }
        apply_modifier(&pair->model, compute, invR, invR2, &Eij, &Wij, 1);
        if (compute) *Epair = *Epair + Eij;
        *Wpair = *Wpair + Wij;
        Wsum = Wij;
        if (ijcharged && pair->coulomb) {
          /* The generated virial dispatch has no "coul_none" case
             (make_virial_compute.sh:26-30), so in virial-only mode Wij then
             keeps the pair-model value. */
          if (compute) coul_compute(cm, &Eij, &Wij, invR, invR2);
          else if (cm->kind != K_COUL_NONE) Wij = coul_virial(cm, invR, invR2);
          apply_modifier(cm, compute, invR, invR2, &Eij, &Wij, 0);
          double QiQj = pair->kCoul * Qi * me->charge[j];
          if (compute) *Ecoul = *Ecoul + QiQj * Eij;
          Wij = QiQj * Wij;
          *Wcoul = *Wcoul + Wij;
          Wsum = Wsum + Wij;
        }
        for (int x = 0; x < 3; x++) {
          double Fij = Wsum * invR2 * Rij[x];
          Fi[x] += Fij;
          F[3 * (j - 1) + x] -= Fij;
        }
      }
    }
    for (int x = 0; x < 3; x++) F[3 * (i - 1) + x] += Fi[x];
  }
  double L = *me->Lbox;
  for (int k = 0; k < 3 * N; k++) F[k] = L * F[k];
}

/* kspaceModelClass discount: modelClass_kspace.f90:321-334 */
static void ks_discount(const okspace_t *ks, double *E, double *W, double rsq, double QiQj) {
  double r = sqrt(rsq), x = ks->alpha * r, expmx2 = exp(-x * x);
  *E = -QiQj * (ONE - uerfc(x, expmx2)) / r;
  *W = *E + QiQj * ks->beta * expmx2;
}

/* EmDeeData.f90:443-487 */
static void compute_bonds(odata_t *me, int thread, const double *R, double *F, double *Potential, double *Virial, double *Ecoul) {
  int bonded = me->bonded[me->layer];
  int kspace = me->coul[me->layer - 1].requires_kspace;
  if (!(me->nbonds > 0 && (bonded || kspace))) return;
  int nb = (me->nbonds + me->nthreads - 1) / me->nthreads;
  double L = *me->Lbox, invL2 = ONE / (L * L);
  int mlast = me->nbonds < thread * nb ? me->nbonds : thread * nb;
  for (int m = (thread - 1) * nb + 1; m <= mlast; m++) {
    const ostruct_t *bd = &me->bonds[m - 1];
    double Rij[3], E = 0, W = 0;
    for (int x = 0; x < 3; x++) { Rij[x] = R[3 * (bd->i - 1) + x] - R[3 * (bd->j - 1) + x]; Rij[x] = Rij[x] - anint(Rij[x]); }
    double invR2 = invL2 / (Rij[0] * Rij[0] + Rij[1] * Rij[1] + Rij[2] * Rij[2]);
    if (bonded) {
      if (bd->model.kind == K_BOND_HARMONIC) {   /* bond_harmonic.f90:69-81 */
        double r = ONE / sqrt(invR2), delta = r - bd->model.r0;
        E = bd->model.half_k * delta * delta;
        W = bd->model.minus_k * delta * r;
      }
      *Potential += E;
      *Virial += W;
    } else W = ZERO;
    if (kspace) {
      int it = me->atomType[bd->i], jt = me->atomType[bd->j];
      double QiQj = PAIR(me, it, jt, me->layer).kCoul * me->charge[bd->i] * me->charge[bd->j], EL, WL;
      ks_discount(me->kspace, &EL, &WL, ONE / invR2, QiQj);
      *Ecoul += EL;
      W = W + WL;
    }
    for (int x = 0; x < 3; x++) {
      double Fij = W * invR2 * L * Rij[x];
      F[3 * (bd->i - 1) + x] += Fij;
      F[3 * (bd->j - 1) + x] -= Fij;
    }
  }
}

static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* EmDeeData.f90:554-640, which the reference declares but never calls from
   its force driver; here it is called after the angles, for layers with
   bonded terms.  R is scaled (R/L).  One deliberate difference: atom j
   receives -(Fi + Fk + Fl), so that the four forces sum to zero (the
   reference line :604 adds +(Fi + Fk + Fl)).  The scaled 1-4 pair term of
   :607-630 is skipped when factor14 = 0, the only value a model here has. */
static void compute_dihedrals(odata_t *me, int thread, const double *R, double *F, double *Potential, double *Virial) {
  if (!(me->ndihedrals > 0 && me->bonded[me->layer])) return;
  double L = *me->Lbox;
  int nd = (me->ndihedrals + me->nthreads - 1) / me->nthreads;
  int mlast = me->ndihedrals < thread * nd ? me->ndihedrals : thread * nd;
  for (int m = (thread - 1) * nd + 1; m <= mlast; m++) {
    const ostruct_t *dh = &me->dihedrals[m - 1];
    double rij[3], rkj[3], rlk[3], x[3], y[3], z[3], u[3], v[3], w[3], c[3], Fi[3], Fk[3], Fl[3];
    for (int q = 0; q < 3; q++) {
      double Rj = R[3 * (dh->j - 1) + q], Rk = R[3 * (dh->k - 1) + q];
      rij[q] = R[3 * (dh->i - 1) + q] - Rj;
      rkj[q] = Rk - Rj;
      rlk[q] = R[3 * (dh->l - 1) + q] - Rk;
      rij[q] = rij[q] - anint(rij[q]);
      rkj[q] = rkj[q] - anint(rkj[q]);
      rlk[q] = rlk[q] - anint(rlk[q]);
    }
    double normRkj = sqrt(dot3(rkj, rkj));
    for (int q = 0; q < 3; q++) z[q] = rkj[q] / normRkj;
    double rz = dot3(rij, z);
    for (int q = 0; q < 3; q++) x[q] = rij[q] - rz * z[q];
    double normX = sqrt(dot3(x, x));
    for (int q = 0; q < 3; q++) x[q] = x[q] / normX;
    cross3(z, x, y);
    double a = dot3(x, rlk), b = dot3(y, rlk);
    double phi = atan2(b, a), Ed = ZERO, Fd = ZERO;
    if (dh->model.kind == K_DIHEDRAL_HARMONIC) {
      double delta = phi - dh->model.r0;
      delta = delta - TWO * PI_ * anint(delta / (TWO * PI_));
      Ed = dh->model.half_k * delta * delta;
      Fd = dh->model.minus_k * delta;
    }
    Fd = Fd / (L * (a * a + b * b));
    cross3(rlk, z, c);
    for (int q = 0; q < 3; q++) u[q] = (a * c[q] - b * rlk[q]) / normX;
    cross3(rlk, x, c);
    double zu = dot3(z, u);
    for (int q = 0; q < 3; q++) v[q] = (a * c[q] + zu * rij[q]) / normRkj;
    double zr = dot3(z, rij);
    for (int q = 0; q < 3; q++) w[q] = v[q] + zr * u[q] / normRkj;
    double uy = dot3(u, y), vx = dot3(v, x), wy = dot3(w, y);
    for (int q = 0; q < 3; q++) {
      Fi[q] = Fd * uy * y[q];
      Fl[q] = Fd * (a * y[q] - b * x[q]);
      Fk[q] = -(Fd * (vx * x[q] + wy * y[q]) + Fl[q]);
    }
    for (int q = 0; q < 3; q++) {
      F[3 * (dh->i - 1) + q] += Fi[q];
      F[3 * (dh->k - 1) + q] += Fk[q];
      F[3 * (dh->l - 1) + q] += Fl[q];
      F[3 * (dh->j - 1) + q] -= (Fi[q] + Fk[q] + Fl[q]);
    }
    *Potential += Ed;
    *Virial += L * ((Fi[0] * rij[0] + Fk[0] * rkj[0] + Fl[0] * (rlk[0] + rkj[0])) +
                    (Fi[1] * rij[1] + Fk[1] * rkj[1] + Fl[1] * (rlk[1] + rkj[1])) +
                    (Fi[2] * rij[2] + Fk[2] * rkj[2] + Fl[2] * (rlk[2] + rkj[2])));
  }
}

/* EmDeeData.f90:491-550 */
static void compute_angles(odata_t *me, int thread, const double *R, double *F, double *Potential, double *Virial, double *Ecoul) {
  int bonded = me->bonded[me->layer];
  int kspace = me->coul[me->layer - 1].requires_kspace;
  if (!(me->nangles > 0 && (bonded || kspace))) return;
  int na = (me->nangles + me->nthreads - 1) / me->nthreads;
  double L = *me->Lbox;
  int mlast = me->nangles < thread * na ? me->nangles : thread * na;
  for (int m = (thread - 1) * na + 1; m <= mlast; m++) {
    const ostruct_t *an = &me->angles[m - 1];
    int i = an->i, j = an->j, k = an->k;
    double avec[3], bvec[3];
    for (int x = 0; x < 3; x++) {
      avec[x] = R[3 * (i - 1) + x] - R[3 * (j - 1) + x];
      bvec[x] = R[3 * (k - 1) + x] - R[3 * (j - 1) + x];
      avec[x] = L * (avec[x] - anint(avec[x]));
      bvec[x] = L * (bvec[x] - anint(bvec[x]));
    }
    if (bonded) {
      double aa = avec[0] * avec[0] + avec[1] * avec[1] + avec[2] * avec[2];
      double bb = bvec[0] * bvec[0] + bvec[1] * bvec[1] + bvec[2] * bvec[2];
      double ab = avec[0] * bvec[0] + avec[1] * bvec[1] + avec[2] * bvec[2];
      double theta = acos(ab / sqrt(aa * bb)), Ea = 0, Fa = 0;
      if (an->model.kind == K_ANGLE_HARMONIC) {  /* angle_harmonic.f90:69-79 */
        double delta = theta - an->model.r0;
        Ea = an->model.half_k * delta * delta;
        Fa = an->model.minus_k * delta;
      }
      double factor = Fa / sqrt(aa * bb - ab * ab), Fi[3], Fk[3];
      for (int x = 0; x < 3; x++) {
        Fi[x] = ((ab / aa) * avec[x] - bvec[x]) * factor;
        Fk[x] = ((ab / bb) * bvec[x] - avec[x]) * factor;
      }
      for (int x = 0; x < 3; x++) {
        F[3 * (i - 1) + x] += Fi[x];
        F[3 * (k - 1) + x] += Fk[x];
        F[3 * (j - 1) + x] -= (Fi[x] + Fk[x]);
      }
      *Potential += Ea;
      *Virial += (Fi[0] * avec[0] + Fk[0] * bvec[0]) + (Fi[1] * avec[1] + Fk[1] * bvec[1]) + (Fi[2] * avec[2] + Fk[2] * bvec[2]);
    }
    if (kspace) {
      double Rik[3], RikSq, EL, WL;
      for (int x = 0; x < 3; x++) Rik[x] = avec[x] - bvec[x];
      RikSq = Rik[0] * Rik[0] + Rik[1] * Rik[1] + Rik[2] * Rik[2];
      int it = me->atomType[i], kt = me->atomType[k];
      double QiQk = PAIR(me, it, kt, me->layer).kCoul * me->charge[i] * me->charge[k];
      ks_discount(me->kspace, &EL, &WL, RikSq, QiQk);
      *Ecoul += EL;
      for (int x = 0; x < 3; x++) {
        double Fik = WL * Rik[x] / RikSq;
        F[3 * (i - 1) + x] += Fik;
        F[3 * (k - 1) + x] -= Fik;
      }
    }
  }
}

/* ------------------------------------------------------------------ Ewald */

/* kspace_ewald.f90:82-99 */
static void ewald_set_parameters(okspace_t *ks, double Rc) {
  double s = sqrt(inverse_of_x_plus_ln_x(-log(ks->base.accuracy)));
  ks->alpha = s / Rc;
  ks->kmax = TWO * ks->alpha * s;
}

/* kspace_ewald.f90:103-184 */
static void ewald_update(okspace_t *ks, double L) {
  double unit = TWO * PI_ / L, unitSq = unit * unit;
  int nmax = (int)ceil(ks->kmax / unit);
  if (nmax != ks->nmax[0] || nmax != ks->nmax[1] || nmax != ks->nmax[2]) {
    ks->nmax[0] = ks->nmax[1] = ks->nmax[2] = nmax;
    int M = 2 * nmax + 1;
    int maxnvecs = M * M * M / 2, M2 = M, M2M3 = M * M, nvecs = 0;
    int *n = xcalloc(3 * (size_t)maxnvecs, sizeof(int));
    double kmaxSq = unitSq * nmax * nmax;
    for (int i = maxnvecs + 1; i <= 2 * maxnvecs; i++) {
      int n1 = i / M2M3, j = i - n1 * M2M3, n2 = j / M2, n3 = j - n2 * M2;
      int kv[3] = { n1 - nmax, n2 - nmax, n3 - nmax };
      if (unitSq * kv[0] * kv[0] + unitSq * kv[1] * kv[1] + unitSq * kv[2] * kv[2] <= kmaxSq) {
        memcpy(n + 3 * nvecs, kv, sizeof kv);
        nvecs++;
      }
    }
    ks->nvecs = nvecs;
    free(ks->prefac); free(ks->n); free(ks->vec); free(ks->qeikr); free(ks->S);
    ks->prefac = xcalloc(nvecs, sizeof(double));
    ks->n = xcalloc(3 * (size_t)nvecs, sizeof(int));
    ks->vec = xcalloc(3 * (size_t)nvecs, sizeof(double));
    ks->qeikr = xcalloc(2 * (size_t)nvecs * ks->ncharged, sizeof(double));
    ks->S = xcalloc(2 * (size_t)nvecs * ks->ntypes, sizeof(double));
    double B = -0.25 / (ks->alpha * ks->alpha);
    ks->Volume = L * L * L;
    double fourPiByV = 4.0 * PI_ / ks->Volume;
    for (int i = 0; i < nvecs; i++) {
      double k[3], ksq;
      for (int x = 0; x < 3; x++) { ks->n[i + (size_t)nvecs * x] = n[3 * i + x]; k[x] = unit * n[3 * i + x]; }
      ksq = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
      ks->prefac[i] = fourPiByV * exp(ksq * B) / ksq;
      for (int x = 0; x < 3; x++) ks->vec[3 * i + x] = TWO * ks->prefac[i] * k[x];
    }
    free(n);
  } else {
    double factor = ks->Volume / (L * L * L);
    for (int i = 0; i < ks->nvecs; i++) { ks->prefac[i] *= factor; for (int x = 0; x < 3; x++) ks->vec[3 * i + x] *= factor; }
    ks->Volume = L * L * L;
  }
}

/* modelClass_kspace.f90:111-261 */
static void kspace_initialize(odata_t *me, okspace_t *ks, double Rc, double L) {
  const char *task = "kspace model initialization";
  int N = me->natoms;
  int ncharged = 0;
  int *charged = xcalloc(N + 1, sizeof(int));
  for (int i = 1; i <= N; i++) { charged[i] = fabs(me->charge[i]) > EPS_ONE; ncharged += charged[i]; }
  if (ncharged == 0) oerror(task, "system has no charged atoms");
  int nChargedTypes = 0;
  for (int i = 1; i <= N; i++) if (charged[i] && me->atomType[i] > nChargedTypes) nChargedTypes = me->atomType[i];
  int *distinct = xcalloc(nChargedTypes + 1, sizeof(int));
  ks->ntypes = 0;
  for (int i = 1; i <= N; i++) if (charged[i]) {
    int t = me->atomType[i], found = 0;
    for (int k = 0; k < ks->ntypes; k++) if (distinct[k] == t) found = 1;
    if (!found) distinct[ks->ntypes++] = t;
  }
  /* sorted ascending (global.f90:68-97) */
  for (int a = 0; a < ks->ntypes; a++) for (int b = a + 1; b < ks->ntypes; b++)
    if (distinct[b] < distinct[a]) { int t = distinct[a]; distinct[a] = distinct[b]; distinct[b] = t; }
  ks->type = distinct;
  ks->ncharged = ncharged;
  ks->cfirst = xcalloc(ks->ntypes + 1, sizeof(int));
  ks->clast = xcalloc(ks->ntypes + 1, sizeof(int));
  ks->citem = xcalloc(ncharged + 1, sizeof(int));
  ks->cvalue = xcalloc(ncharged + 1, sizeof(double));
  int last = 0;
  for (int t = 1; t <= ks->ntypes; t++) {
    int first = last + 1;
    for (int i = 1; i <= N; i++) if (me->atomType[i] == ks->type[t - 1]) { last++; ks->citem[last] = i; ks->cvalue[last] = me->charge[i]; }
    ks->cfirst[t] = first; ks->clast[t] = last;
  }
  ewald_set_parameters(ks, Rc);
  ks->beta = TWO * ks->alpha / sqrt(PI_);
  /* rigid intrabody pairs and self energy */
  ks->nTypePairs = ks->ntypes * (ks->ntypes - 1) / 2 + ks->ntypes;
  ks->Erigid = xcalloc(ks->nTypePairs + 1, sizeof(double));
  int maxnpairs = 0;
  for (int b = 1; b <= me->nbodies; b++) maxnpairs += me->body[b].NP * (me->body[b].NP - 1) / 2;
  ks->rp_i = xcalloc(maxnpairs + 1, sizeof(int)); ks->rp_j = xcalloc(maxnpairs + 1, sizeof(int));
  ks->rp_it = xcalloc(maxnpairs + 1, sizeof(int)); ks->rp_jt = xcalloc(maxnpairs + 1, sizeof(int));
  ks->rp_FbyR = xcalloc(maxnpairs + 1, sizeof(double));
  int *local = xcalloc(nChargedTypes + 1, sizeof(int));
  for (int t = 1; t <= ks->ntypes; t++) local[ks->type[t - 1]] = t;
  int npairs = 0;
  double invL = ONE / L;
  for (int b = 1; b <= me->nbodies; b++) {
    obody_t *bd = &me->body[b];
    int n = bd->NP;
    for (int ii = 1; ii <= n - 1; ii++) {
      int i = bd->index[ii - 1];
      if (!charged[i]) continue;
      const double *Ri = me->R + 3 * (i - 1);
      double Qi = me->charge[i];
      int itype = local[me->atomType[i]];
      for (int jj = ii + 1; jj <= n; jj++) {
        int j = bd->index[jj - 1];
        if (!charged[j]) continue;
        int jtype = local[me->atomType[j]];
        double Rij[3], rsq, Eij, Wij;
        for (int x = 0; x < 3; x++) { Rij[x] = Ri[x] - me->R[3 * (j - 1) + x]; Rij[x] = Rij[x] - L * anint(invL * Rij[x]); }
        rsq = Rij[0] * Rij[0] + Rij[1] * Rij[1] + Rij[2] * Rij[2];
        ks_discount(ks, &Eij, &Wij, rsq, Qi * me->charge[j]);
        int k = symm1D(itype, jtype);
        ks->Erigid[k] += Eij;
        npairs++;
        ks->rp_i[npairs] = i; ks->rp_j[npairs] = j; ks->rp_it[npairs] = itype; ks->rp_jt[npairs] = jtype;
        ks->rp_FbyR[npairs] = Wij / rsq;
      }
    }
  }
  for (int t = 1; t <= ks->ntypes; t++) {
    int k = symm1D(t, t);
    double s = 0;
    for (int p = ks->cfirst[t]; p <= ks->clast[t]; p++) s += ks->cvalue[p] * ks->cvalue[p];
    ks->Erigid[k] = ks->Erigid[k] - ks->alpha * s / sqrt(PI_);
  }
  ks->nRigidPairs = npairs;
  ks->nlayers = me->nlayers;
  ks->lambda = xcalloc((size_t)ks->ntypes * ks->ntypes * ks->nlayers, sizeof(double));
  ks->lambda1D = xcalloc((size_t)(ks->nTypePairs + 1) * ks->nlayers, sizeof(double));
  for (int i = 1; i <= ks->ntypes; i++) {
    int itype = ks->type[i - 1];
    for (int j = i; j <= ks->ntypes; j++) {
      int jtype = ks->type[j - 1];
      for (int l = 1; l <= ks->nlayers; l++) {
        double lam = PAIR(me, itype, jtype, l).kCoul;
        ks->lambda[(i - 1) + ks->ntypes * ((j - 1) + ks->ntypes * (l - 1))] = lam;
        ks->lambda[(j - 1) + ks->ntypes * ((i - 1) + ks->ntypes * (l - 1))] = lam;
        ks->lambda1D[symm1D(i, j) + (size_t)(ks->nTypePairs + 1) * (l - 1)] = lam;
      }
    }
  }
  ks->nthreads = me->nthreads;
  ewald_update(ks, L);
  free(charged); free(local);
}

/* kspace_ewald.f90:188-248 */
static void ewald_prepare(okspace_t *ks, const double *Rs) {
  int nv = ks->nvecs, nm = ks->nmax[0];
  double *G1 = xcalloc(2 * (nm + 1), sizeof(double));
  double *G2 = xcalloc(2 * (2 * nm + 1), sizeof(double));
  double *G3 = xcalloc(2 * (2 * nm + 1), sizeof(double));
  double *g2 = G2 + 2 * nm, *g3 = G3 + 2 * nm;   /* index range -nm..nm */
  G1[0] = 1; G1[1] = 0; g2[0] = 1; g2[1] = 0; g3[0] = 1; g3[1] = 0;
  for (int j = 1; j <= ks->ncharged; j++) {
    int i = ks->citem[j];
    double Z[3][2];
    for (int x = 0; x < 3; x++) { double th = TWO * PI_ * Rs[3 * (i - 1) + x]; Z[x][0] = cos(th); Z[x][1] = sin(th); }
    double *G[3] = { G1, g2, g3 };
    for (int x = 0; x < 3; x++) {      /* recursion: G(j) = Z**j by repeated products */
      double br = Z[x][0], bi = Z[x][1];
      for (int p = 1; p <= nm; p++) {
        G[x][2 * p] = br; G[x][2 * p + 1] = bi;
        double nr = br * Z[x][0] - bi * Z[x][1], ni = br * Z[x][1] + bi * Z[x][0];
        br = nr; bi = ni;
      }
    }
    for (int p = 1; p <= nm; p++) {
      g2[-2 * p] = g2[2 * p]; g2[-2 * p + 1] = -g2[2 * p + 1];
      g3[-2 * p] = g3[2 * p]; g3[-2 * p + 1] = -g3[2 * p + 1];
    }
    double q = ks->cvalue[j];
    for (int v = 0; v < nv; v++) {
      int a = ks->n[v], b = ks->n[v + nv], c = ks->n[v + 2 * nv];
      double r = q * G1[2 * a], im = q * G1[2 * a + 1];
      double r2 = r * g2[2 * b] - im * g2[2 * b + 1], i2 = r * g2[2 * b + 1] + im * g2[2 * b];
      double r3 = r2 * g3[2 * c] - i2 * g3[2 * c + 1], i3 = r2 * g3[2 * c + 1] + i2 * g3[2 * c];
      ks->qeikr[2 * (v + (size_t)nv * (j - 1))] = r3;
      ks->qeikr[2 * (v + (size_t)nv * (j - 1)) + 1] = i3;
    }
  }
  for (int t = 1; t <= ks->ntypes; t++)
    for (int v = 0; v < nv; v++) {
      double sr = 0, si = 0;
      for (int j = ks->cfirst[t]; j <= ks->clast[t]; j++) {
        sr += ks->qeikr[2 * (v + (size_t)nv * (j - 1))];
        si += ks->qeikr[2 * (v + (size_t)nv * (j - 1)) + 1];
      }
      ks->S[2 * (v + (size_t)nv * (t - 1))] = sr;
      ks->S[2 * (v + (size_t)nv * (t - 1)) + 1] = si;
    }
  free(G1); free(G2); free(G3);
}

/* kspace_ewald.f90:252-314 */
static void ewald_compute(okspace_t *ks, int layer, double *E, double *F) {
  int nv = ks->nvecs, nt = ks->ntypes;
  double *sigma = xcalloc(2 * (size_t)nv * nt, sizeof(double));
  const double *lam = ks->lambda + (size_t)nt * nt * (layer - 1);
  double En = 0;
  for (int v = 0; v < nv; v++) {
    double s = 0;
    for (int t2 = 0; t2 < nt; t2++) {
      double sr = 0, si = 0;
      for (int t1 = 0; t1 < nt; t1++) {
        sr += ks->S[2 * (v + (size_t)nv * t1)] * lam[t1 + nt * t2];
        si += ks->S[2 * (v + (size_t)nv * t1) + 1] * lam[t1 + nt * t2];
      }
      sigma[2 * (v + (size_t)nv * t2)] = sr;
      sigma[2 * (v + (size_t)nv * t2) + 1] = si;
      s += ks->S[2 * (v + (size_t)nv * t2)] * sr + ks->S[2 * (v + (size_t)nv * t2) + 1] * si;
    }
    En += ks->prefac[v] * s;
  }
  *E += En;
  int itype = 1;
  for (int j = 1; j <= ks->ncharged; j++) {
    int i = ks->citem[j];
    if (j > ks->clast[itype]) itype++;
    double f[3] = { 0, 0, 0 };
    for (int v = 0; v < nv; v++) {
      double sr = sigma[2 * (v + (size_t)nv * (itype - 1))], si = sigma[2 * (v + (size_t)nv * (itype - 1)) + 1];
      double qr = ks->qeikr[2 * (v + (size_t)nv * (j - 1))], qi = ks->qeikr[2 * (v + (size_t)nv * (j - 1)) + 1];
      double c = sr * qi - qr * si;
      for (int x = 0; x < 3; x++) f[x] += ks->vec[3 * v + x] * c;
    }
    for (int x = 0; x < 3; x++) F[3 * (i - 1) + x] += f[x];
  }
  free(sigma);
}

/* EmDeeData.f90:689-700 */
static void compute_kspace(odata_t *me, const double *Rs, double *Elong, double *F) {
  okspace_t *ks = me->kspace;
  ewald_prepare(ks, Rs);
  ewald_compute(ks, me->layer, Elong, F);
  /* discount_rigid_pairs, energy only: modelClass_kspace.f90:265-317 */
  double s = 0;
  for (int k = 1; k <= ks->nTypePairs; k++) s += ks->lambda1D[k + (size_t)(ks->nTypePairs + 1) * (me->layer - 1)] * ks->Erigid[k];
  *Elong += s;
}

/* EmDeeData.f90:926-953 */
static double rigid_body_virial(odata_t *me) {
  double W = 0;
  for (int b = 1; b <= me->nbodies; b++) {
    obody_t *bd = &me->body[b];
    for (int k = 0; k < bd->NP; k++)
      for (int x = 0; x < 3; x++) W += me->F[3 * (bd->index[k] - 1) + x] * bd->delta[3 * k + x];
  }
  return -W;
}

/* EmDeeCode.f90:1215-1277 */
void EmDee_compute_forces(tEmDee *md) {
  odata_t *me = ME(*md);
  int N = me->natoms, T = me->nthreads;
  double E[6] = { 0 }, W[6] = { 0 };
  enum { PAIR_ = 0, COUL_ = 1, LONG_ = 2, BOND_ = 3, ANGLE_ = 4, DIHEDRAL_ = 5 };
  double *Rs = xcalloc(3 * (size_t)N, sizeof(double));
  double *Fs = xcalloc(3 * (size_t)N * T, sizeof(double));
  double L = *me->Lbox;
  for (int k = 0; k < 3 * N; k++) Rs[k] = me->R[k] / L;
  handle_neighbor_lists(me, &md->Builds, &md->Time.Neighbor, Rs);
  md->Time.Pair -= wtime();
  int compute = md->Options.Compute;
  double Et[6 * 64] = { 0 }, Wt[6 * 64] = { 0 };
  if (T > 64) oerror("oracle", "at most 64 threads");
#pragma omp parallel num_threads(T)
  {
    int t = thread_num() + 1;
    double *F = Fs + 3 * (size_t)N * (t - 1), *e = Et + 6 * (t - 1), *w = Wt + 6 * (t - 1);
    compute_pairs(me, t, compute, Rs, F, &e[PAIR_], &e[COUL_], &w[PAIR_], &w[COUL_]);
    compute_bonds(me, t, Rs, F, &e[BOND_], &w[BOND_], &e[COUL_]);
    compute_angles(me, t, Rs, F, &e[ANGLE_], &w[ANGLE_], &e[COUL_]);
    compute_dihedrals(me, t, Rs, F, &e[DIHEDRAL_], &w[DIHEDRAL_]);
  }
  for (int t = 0; t < T; t++) for (int k = 0; k < 6; k++) { E[k] += Et[6 * t + k]; W[k] += Wt[6 * t + k]; }
  for (int k = 0; k < 3 * N; k++) {
    double s = 0;
    for (int t = 0; t < T; t++) s += Fs[3 * (size_t)N * t + k];
    me->F[k] = s;
  }
  if (me->coul[me->layer - 1].requires_kspace) {
    compute_kspace(me, Rs, &E[LONG_], me->F);
    W[LONG_] = E[COUL_] + E[LONG_] - W[COUL_];
  }
  md->Virial = W[0] + W[1] + W[2] + W[3] + W[4] + W[5];
  if (me->nbodies != 0) {
    md->BodyVirial = rigid_body_virial(me);
    md->Virial = md->Virial + md->BodyVirial;
  }
  if (compute) {
    md->Energy.Dispersion = E[PAIR_];
    md->Energy.Coulomb = E[COUL_] + E[LONG_];
    md->Energy.Bond = E[BOND_];
    md->Energy.Angle = E[ANGLE_];
    md->Energy.Dihedral = E[DIHEDRAL_];
    md->Energy.Potential = E[0] + E[1] + E[2] + E[3] + E[4] + E[5];
    md->Energy.ShadowPotential = md->Energy.Potential;
  }
  md->Energy.UpToDate = compute;
  me->forcesUpToDate[me->layer] = 1;
  memcpy(layer_energy_ptr(me, me->layer), &md->Energy, sizeof(oenergy_t));
  ((ovirial_t *)me->layerVirial)[me->layer - 1].Total = md->Virial;
  ((ovirial_t *)me->layerVirial)[me->layer - 1].Body = md->BodyVirial;
  double t = wtime();
  md->Time.Pair += t;
  md->Time.Total = t - me->startTime;
  free(Rs); free(Fs);
}

/* ============================================================ motion */

/* EmDeeData.f90:420-439 */
static void update_rigid_bodies(odata_t *me) {
  double L = *me->Lbox, invL = ONE / L;
  for (int j = 1; j <= me->nbodies; j++) {
    obody_t *b = &me->body[j];
    double *R = xcalloc(3 * (size_t)b->NP, sizeof(double));
    for (int k = 0; k < b->NP; k++) for (int x = 0; x < 3; x++) R[3 * k + x] = me->R[3 * (b->index[k] - 1) + x];
    for (int k = 1; k < b->NP; k++)
      for (int x = 0; x < 3; x++) R[3 * k + x] = R[3 * k + x] - L * anint(invL * (R[3 * k + x] - R[x]));
    body_update(b, R);
    for (int k = 0; k < b->NP; k++) for (int x = 0; x < 3; x++) me->R[3 * (b->index[k] - 1) + x] = R[3 * k + x];
    free(R);
  }
}

/* EmDeeData.f90:194-223 */
static void check_actual_interactions(odata_t *me) {
  int nt = me->ntypes, nl = me->nlayers;
  int *neutral = xcalloc(nt + 1, sizeof(int));
  int *inter = xcalloc((size_t)nt * nt * nl, sizeof(int));
  for (int i = 1; i <= nt; i++) {
    int c = 0;
    for (int a = 1; a <= me->natoms; a++) if (me->atomType[a] == i && me->charged[a]) c++;
    neutral[i] = c == 0;
  }
  for (int i = 1; i <= nt; i++)
    for (int j = 1; j <= i; j++) {
      int any = 0;
      for (int k = 1; k <= nl; k++) {
        const opair_t *p = &PAIR(me, i, j, k);
        int no_pair = p->model.kind == K_PAIR_NONE;
        int no_coul = me->coul[k - 1].kind == K_COUL_NONE || !p->coulomb;
        int coul_only = no_pair && !no_coul;
        int inert = (no_pair && no_coul) || (coul_only && neutral[i] && neutral[j]);
        inter[(i - 1) + nt * ((j - 1) + nt * (k - 1))] = !inert;
        any |= !inert;
      }
      T2(me, interact, i, j) = any;
      T2(me, interact, j, i) = any;
    }
  for (int k = 1; k <= nl; k++) {
    me->pairs_exist[k] = 0;
    for (int i = 1; i <= nt; i++) for (int j = 1; j <= i; j++) me->pairs_exist[k] |= inter[(i - 1) + nt * ((j - 1) + nt * (k - 1))];
  }
  free(neutral); free(inter);
}

/* EmDeeData.f90:359-416 */
static void perform_initialization(odata_t *me, int *DoF, int *RotDoF) {
  const char *task = "system initialization";
  update_rigid_bodies(me);
  int bodyDoF = 0;
  for (int b = 1; b <= me->nbodies; b++) bodyDoF += me->body[b].dof;
  *RotDoF = bodyDoF - 3 * me->nbodies;
  *DoF = 3 * me->nfree + bodyDoF - 3;
  check_actual_interactions(me);
  int any_required = 0, first_req = 0;
  double kspaceRc = 0;
  for (int l = 1; l <= me->nlayers; l++)
    if (me->coul[l - 1].requires_kspace) {
      double lrc = me->useInRc[l] ? me->InRc : me->Rc;
      if (!any_required) { kspaceRc = lrc; first_req = l; }
      else if (lrc != kspaceRc) oerror(task, "all layers with ewald-like coulomb models must have the same cutoff");
      any_required = 1;
    }
  (void)first_req;
  if (any_required != me->kspace_active) {
    if (me->kspace_active) me->kspace_active = 0;
    else oerror(task, "a kspace solver is required, but has not been defined");
  }
  if (me->kspace_active) {
    kspace_initialize(me, me->kspace, kspaceRc, *me->Lbox);
    for (int l = 1; l <= me->nlayers; l++)
      if (me->coul[l - 1].requires_kspace) coul_kspace_setup(&me->coul[l - 1], me->kspace->alpha);
  }
  me->initialized = 1;
}

/* EmDeeData.f90:157-189 */
static void assign_momenta(odata_t *me, const double *P, double *twoKEt, double *twoKEr) {
  for (int x = 0; x < 3; x++) { twoKEt[x] = 0; twoKEr[x] = 0; }
  for (int j = 1; j <= me->nfree; j++) {
    int i = me->free_[j];
    for (int x = 0; x < 3; x++) {
      me->P[3 * (i - 1) + x] = P[3 * (i - 1) + x];
      twoKEt[x] += me->invMass[i] * (P[3 * (i - 1) + x] * P[3 * (i - 1) + x]);
    }
  }
  for (int ib = 1; ib <= me->nbodies; ib++) {
    obody_t *b = &me->body[ib];
    double Lv[3] = { 0, 0, 0 };
    for (int x = 0; x < 3; x++) b->pcm[x] = 0;
    for (int j = 0; j < b->NP; j++) {
      const double *Pj = P + 3 * (b->index[j] - 1);
      double cr[3];
      for (int x = 0; x < 3; x++) b->pcm[x] += Pj[x];
      cross3(b->delta + 3 * j, Pj, cr);
      for (int x = 0; x < 3; x++) Lv[x] += cr[x];
    }
    for (int x = 0; x < 3; x++) twoKEt[x] += b->invMass * (b->pcm[x] * b->pcm[x]);
    double C[12], B[12], t[3] = { TWO * Lv[0], TWO * Lv[1], TWO * Lv[2] }, s[3];
    matC(b->q, C);
    mul43(C, t, b->pi);
    matB(b->q, B);
    mul34t(B, b->pi, s);
    for (int x = 0; x < 3; x++) b->omega[x] = HALF * b->invMoI[x] * s[x];
    for (int x = 0; x < 3; x++) twoKEr[x] += b->MoI[x] * (b->omega[x] * b->omega[x]);
  }
}

static void switch_layer_internal(tEmDee *md, int layer);

/* EmDeeCode.f90:914-923 */
static void initialize_system(tEmDee *md) {
  odata_t *me = ME(*md);
  perform_initialization(me, &md->DoF, &md->RotDoF);
  for (int layer = me->nlayers; layer >= 1; layer--) {
    switch_layer_internal(md, layer);
    EmDee_compute_forces(md);
  }
}

static void mark_stale(odata_t *me, tEmDee *md) {
  for (int l = 1; l <= me->nlayers; l++) { me->forcesUpToDate[l] = 0; ((oenergy_t *)me->layerEnergy)[l - 1].UpToDate = 0; }
  md->Energy.UpToDate = 0;
}

/* EmDeeCode.f90:805-925 */
void EmDee_upload(tEmDee *md, char *option, double *address) {
  odata_t *me = ME(*md);
  int N = me->natoms;
  if (!address) oerror("upload", "provided address is invalid");
  if (!strcmp(option, "box")) {
    if (!me->Lbox) me->Lbox = xcalloc(1, sizeof(double));
    *me->Lbox = *address;
    if (me->initialized) mark_stale(me, md);
    else if (me->R) initialize_system(md);
  } else if (!strcmp(option, "coordinates")) {
    if (!me->R) me->R = xcalloc(3 * (size_t)N, sizeof(double));
    memcpy(me->R, address, 3 * sizeof(double) * N);
    if (me->initialized) {
      mark_stale(me, md);
      if (md->Options.AutoBodyUpdate) update_rigid_bodies(me);
    } else if (me->Lbox) initialize_system(md);
  } else if (!strcmp(option, "momenta")) {
    if (!me->initialized) oerror("upload", "box and coordinates have not been defined");
    double t2[3], r2[3];
    assign_momenta(me, address, t2, r2);
    for (int x = 0; x < 3; x++) { md->Kinetic.RotPart[x] = HALF * r2[x]; md->Kinetic.TransPart[x] = HALF * t2[x]; }
    md->Kinetic.Rotational = md->Kinetic.RotPart[0] + md->Kinetic.RotPart[1] + md->Kinetic.RotPart[2];
    md->Kinetic.Kinetic = (md->Kinetic.TransPart[0] + md->Kinetic.TransPart[1] + md->Kinetic.TransPart[2]) + md->Kinetic.Rotational;
    md->Kinetic.ShadowKinetic = md->Kinetic.Kinetic;
    md->Kinetic.ShadowRotational = md->Kinetic.Rotational;
    md->Kinetic.UpToDate = 1;
  } else if (!strcmp(option, "forces")) {
    if (!me->initialized) oerror("upload", "box and coordinates have not been defined");
    memcpy(me->F, address, 3 * sizeof(double) * N);
  } else if (!strcmp(option, "charges")) {
    if (me->initialized) oerror("upload", "cannot set charges after box and coordinates initialization");
    for (int i = 1; i <= N; i++) { me->charge[i] = address[i - 1]; me->charged[i] = fabs(address[i - 1]) > EPS_ONE; }
    mark_stale(me, md);
  } else oerror("upload", "invalid option");
}

/* EmDeeCode.f90:659-801 */
void EmDee_download(tEmDee md, char *option, double *address) {
  odata_t *me = ME(md);
  int N = me->natoms;
  if (!address) oerror("download", "provided address is invalid");
  if (!strcmp(option, "box")) *address = *me->Lbox;
  else if (!strcmp(option, "coordinates")) {
    if (!me->R) oerror("download", "coordinates have not been allocated");
    memcpy(address, me->R, 3 * sizeof(double) * N);
  } else if (!strcmp(option, "momenta")) {
    for (int j = 1; j <= me->nfree; j++) { int i = me->free_[j]; for (int x = 0; x < 3; x++) address[3 * (i - 1) + x] = me->P[3 * (i - 1) + x]; }
    for (int b = 1; b <= me->nbodies; b++) {
      obody_t *bd = &me->body[b];
      double *P = xcalloc(3 * (size_t)bd->NP, sizeof(double));
      body_particle_momenta(bd, P);
      for (int k = 0; k < bd->NP; k++) for (int x = 0; x < 3; x++) address[3 * (bd->index[k] - 1) + x] = P[3 * k + x];
      free(P);
    }
  } else if (!strcmp(option, "forces")) {
    if (!me->forcesUpToDate[me->layer]) EmDee_compute_forces(&md);
    memcpy(address, me->F, 3 * sizeof(double) * N);
  } else if (!strcmp(option, "centersOfMass")) {
    for (int b = 1; b <= me->nbodies; b++) for (int x = 0; x < 3; x++) address[3 * (b - 1) + x] = me->body[b].rcm[x];
    for (int j = 1; j <= me->nfree; j++) for (int x = 0; x < 3; x++) address[3 * (j + me->nbodies - 1) + x] = me->R[3 * (me->free_[j] - 1) + x];
  } else if (!strcmp(option, "quaternions") || !strcmp(option, "quatmom") || !strcmp(option, "quattau")) {
    for (int b = 1; b <= me->nbodies; b++) {
      obody_t *bd = &me->body[b];
      double *o = address + 4 * (b - 1);
      if (!strcmp(option, "quaternions")) memcpy(o, bd->q, 4 * sizeof(double));
      else if (!strcmp(option, "quatmom")) memcpy(o, bd->pi, 4 * sizeof(double));
      else { double C[12], t[3] = { TWO * bd->tau[0], TWO * bd->tau[1], TWO * bd->tau[2] }; matC(bd->q, C); mul43(C, t, o); }
    }
  } else if (!strcmp(option, "angmom") || !strcmp(option, "bodycoord") || !strcmp(option, "bodymom") ||
             !strcmp(option, "bodyforces") || !strcmp(option, "torques") || !strcmp(option, "inertia")) {
    for (int b = 1; b <= me->nbodies; b++) {
      obody_t *bd = &me->body[b];
      const double *src = !strcmp(option, "angmom") ? bd->omega : !strcmp(option, "bodycoord") ? bd->rcm :
                          !strcmp(option, "bodymom") ? bd->pcm : !strcmp(option, "bodyforces") ? bd->F :
                          !strcmp(option, "torques") ? bd->tau : bd->MoI;
      memcpy(address + 3 * (b - 1), src, 3 * sizeof(double));
    }
  } else oerror("download", "invalid option");
}

static void switch_layer_internal(tEmDee *md, int layer) {
  odata_t *me = ME(*md);
  if (layer != me->layer) {
    if (layer < 1 || layer > me->nlayers) oerror("model layer switch", "selected layer is out of range");
    me->layer = layer;
    memcpy(&md->Energy, layer_energy_ptr(me, layer), sizeof(oenergy_t));
    md->Virial = ((ovirial_t *)me->layerVirial)[layer - 1].Total;
    md->BodyVirial = ((ovirial_t *)me->layerVirial)[layer - 1].Body;
    me->F = me->layerF + 3 * (size_t)me->natoms * (layer - 1);
  }
}
/* EmDeeCode.f90:929-946 */
void EmDee_switch_model_layer(tEmDee *md, int layer) { switch_layer_internal(md, layer); }

/* EmDeeCode.f90:950-1020 */
void EmDee_random_momenta(tEmDee *md, double kT, bool adjust, int seed) {
  odata_t *me = ME(*md);
  if (me->random.seeding_required) rng_setup(&me->random, seed);
  double twoKEt[3] = { 0, 0, 0 }, twoKEr[3] = { 0, 0, 0 };
  okiss_t *rng = &me->random;
  if (me->nbodies != 0) {
    if (!me->initialized) oerror("random_momenta", "coordinates have not defined");
    for (int i = 1; i <= me->nbodies; i++) {
      obody_t *b = &me->body[i];
      double f = sqrt(b->mass * kT);
      for (int x = 0; x < 3; x++) b->pcm[x] = f * rng_normal(rng);
      double w[3];
      for (int x = 0; x < 3; x++) w[x] = sqrt(b->invMoI[x] * kT) * rng_normal(rng);
      body_assign_momenta(b, w, 3);
      for (int x = 0; x < 3; x++) { twoKEt[x] += b->invMass * (b->pcm[x] * b->pcm[x]); twoKEr[x] += b->MoI[x] * (b->omega[x] * b->omega[x]); }
    }
  }
  for (int j = 1; j <= me->nfree; j++) {
    int i = me->free_[j];
    double f = sqrt(me->mass[i] * kT);
    for (int x = 0; x < 3; x++) me->P[3 * (i - 1) + x] = f * rng_normal(rng);
    for (int x = 0; x < 3; x++) twoKEt[x] += me->invMass[i] * (me->P[3 * (i - 1) + x] * me->P[3 * (i - 1) + x]);
  }
  if (adjust) {   /* adjust_momenta: EmDeeCode.f90:984-1018 */
    double vcm[3];
    for (int x = 0; x < 3; x++) {
      double s1 = 0, s2 = 0;
      for (int j = 1; j <= me->nfree; j++) s1 += me->P[3 * (me->free_[j] - 1) + x];
      for (int b = 1; b <= me->nbodies; b++) s2 += me->body[b].pcm[x];
      vcm[x] = (s1 + s2) / me->totalMass;
    }
    for (int x = 0; x < 3; x++) twoKEt[x] = 0;
    for (int j = 1; j <= me->nfree; j++) {
      int i = me->free_[j];
      for (int x = 0; x < 3; x++) {
        me->P[3 * (i - 1) + x] = me->P[3 * (i - 1) + x] - me->mass[i] * vcm[x];
        twoKEt[x] += me->invMass[i] * (me->P[3 * (i - 1) + x] * me->P[3 * (i - 1) + x]);
      }
    }
    for (int b = 1; b <= me->nbodies; b++) {
      obody_t *bd = &me->body[b];
      for (int x = 0; x < 3; x++) {
        bd->pcm[x] = bd->pcm[x] - bd->mass * vcm[x];
        twoKEt[x] += bd->invMass * (bd->pcm[x] * bd->pcm[x]);
      }
    }
    int dofsum = 0;
    for (int b = 1; b <= me->nbodies; b++) dofsum += me->body[b].dof;
    double tot = (twoKEt[0] + twoKEr[0]) + (twoKEt[1] + twoKEr[1]) + (twoKEt[2] + twoKEr[2]);
    double factor = sqrt((double)(3 * me->nfree + dofsum - 3) * kT / tot);
    for (int j = 1; j <= me->nfree; j++) for (int x = 0; x < 3; x++) me->P[3 * (me->free_[j] - 1) + x] *= factor;
    for (int b = 1; b <= me->nbodies; b++) {
      obody_t *bd = &me->body[b];
      double w[3];
      for (int x = 0; x < 3; x++) { bd->pcm[x] = factor * bd->pcm[x]; w[x] = factor * bd->omega[x]; }
      body_assign_momenta(bd, w, 3);
    }
    for (int x = 0; x < 3; x++) { twoKEt[x] = factor * factor * twoKEt[x]; twoKEr[x] = factor * factor * twoKEr[x]; }
  }
  for (int x = 0; x < 3; x++) { md->Kinetic.RotPart[x] = HALF * twoKEr[x]; md->Kinetic.TransPart[x] = HALF * twoKEt[x]; }
  md->Kinetic.Rotational = md->Kinetic.RotPart[0] + md->Kinetic.RotPart[1] + md->Kinetic.RotPart[2];
  md->Kinetic.Kinetic = (md->Kinetic.TransPart[0] + md->Kinetic.TransPart[1] + md->Kinetic.TransPart[2]) + md->Kinetic.Rotational;
  md->Kinetic.ShadowKinetic = md->Kinetic.Kinetic;
  md->Kinetic.ShadowRotational = md->Kinetic.Rotational;
  md->Kinetic.UpToDate = 1;
}

/* EmDeeData.f90:864-896 */
static void boost_all(odata_t *me, double P_factor, double F_factor, const double *F, int translate, int rotate) {
  for (int i = 1; i <= me->nbodies; i++) body_force_and_torque(&me->body[i], F);
  if (translate) {
    for (int i = 1; i <= me->nbodies; i++)
      for (int x = 0; x < 3; x++) me->body[i].pcm[x] = P_factor * me->body[i].pcm[x] + F_factor * me->body[i].F[x];
    for (int j = 1; j <= me->nfree; j++) {
      int i = me->free_[j];
      for (int x = 0; x < 3; x++) me->P[3 * (i - 1) + x] = P_factor * me->P[3 * (i - 1) + x] + F_factor * F[3 * (i - 1) + x];
    }
  }
  if (rotate) {
    double Ctau = TWO * F_factor;
    for (int i = 1; i <= me->nbodies; i++) {
      obody_t *b = &me->body[i];
      double C[12], t[3] = { Ctau * b->tau[0], Ctau * b->tau[1], Ctau * b->tau[2] }, c4[4], in[4];
      matC(b->q, C);
      mul43(C, t, c4);
      for (int k = 0; k < 4; k++) in[k] = P_factor * b->pi[k] + c4[k];
      body_assign_momenta(b, in, 4);
    }
  }
}

/* EmDeeData.f90:900-922 */
static void kinetic_energies(odata_t *me, int translate, int rotate, double *twoKEt, double *twoKEr) {
  if (translate) {
    for (int x = 0; x < 3; x++) twoKEt[x] = 0;
    for (int i = 1; i <= me->nbodies; i++) for (int x = 0; x < 3; x++) twoKEt[x] += me->body[i].invMass * me->body[i].pcm[x] * me->body[i].pcm[x];
    for (int j = 1; j <= me->nfree; j++) {
      int i = me->free_[j];
      for (int x = 0; x < 3; x++) twoKEt[x] += me->invMass[i] * (me->P[3 * (i - 1) + x] * me->P[3 * (i - 1) + x]);
    }
  }
  if (rotate) {
    for (int x = 0; x < 3; x++) twoKEr[x] = 0;
    for (int i = 1; i <= me->nbodies; i++) for (int x = 0; x < 3; x++) twoKEr[x] += me->body[i].MoI[x] * me->body[i].omega[x] * me->body[i].omega[x];
  }
}

/* EmDeeCode.f90:1024-1065 */
void EmDee_boost(tEmDee *md, double lambda, double alpha, double dt) {
  odata_t *me = ME(*md);
  double CF = phi_fn(alpha * dt) * dt, CP = ONE - alpha * CF;
  CF = lambda * CF;
  if (lambda != ZERO && !me->forcesUpToDate[me->layer]) EmDee_compute_forces(md);
  double twoKEt[3] = { 0, 0, 0 }, twoKEr[3] = { 0, 0, 0 };
  boost_all(me, CP, CF, me->F, md->Options.Translate, md->Options.Rotate);
  if (md->Options.Compute) {
    kinetic_energies(me, md->Options.Translate, md->Options.Rotate, twoKEt, twoKEr);
    if (md->Options.Translate) for (int x = 0; x < 3; x++) md->Kinetic.TransPart[x] = HALF * twoKEt[x];
    if (md->Options.Rotate) {
      for (int x = 0; x < 3; x++) md->Kinetic.RotPart[x] = HALF * twoKEr[x];
      md->Kinetic.Rotational = md->Kinetic.RotPart[0] + md->Kinetic.RotPart[1] + md->Kinetic.RotPart[2];
    }
    md->Kinetic.Kinetic = (md->Kinetic.TransPart[0] + md->Kinetic.TransPart[1] + md->Kinetic.TransPart[2]) + md->Kinetic.Rotational;
  }
  md->Kinetic.UpToDate = md->Options.Compute;
}

/* EmDeeData.f90:823-860 */
static void move_all(odata_t *me, double R_factor, double P_factor, double dt, int translate, int rotate, int mode) {
  if (translate) {
    for (int i = 1; i <= me->nbodies; i++) {
      obody_t *b = &me->body[i];
      for (int x = 0; x < 3; x++) b->rcm[x] = R_factor * b->rcm[x] + P_factor * b->invMass * b->pcm[x];
    }
    for (int j = 1; j <= me->nfree; j++) {
      int i = me->free_[j];
      for (int x = 0; x < 3; x++)
        me->R[3 * (i - 1) + x] = R_factor * me->R[3 * (i - 1) + x] + P_factor * me->P[3 * (i - 1) + x] * me->invMass[i];
    }
  }
  if (rotate) {
    for (int i = 1; i <= me->nbodies; i++) {
      obody_t *b = &me->body[i];
      if (mode == 0) rotate_exact(b, dt); else rotate_no_squish(b, dt, mode);
      for (int k = 0; k < b->NP; k++) for (int x = 0; x < 3; x++) me->R[3 * (b->index[k] - 1) + x] = b->rcm[x] + b->delta[3 * k + x];
    }
  }
}

/* EmDeeCode.f90:1069-1103 */
static void displace_impl(tEmDee *md, double lambda, double alpha, double dt) {
  odata_t *me = ME(*md);
  md->Time.Motion -= wtime();
  double CR, CP;
  if (alpha != ZERO) {
    CP = phi_fn(alpha * dt) * dt;
    CR = ONE - alpha * CP;
    *me->Lbox = CR * *me->Lbox;
  } else { CP = dt; CR = ONE; }
  CP = lambda * CP;
  move_all(me, CR, CP, dt, md->Options.Translate, md->Options.Rotate, md->Options.RotationMode);
  for (int l = 1; l <= me->nlayers; l++) { me->forcesUpToDate[l] = 0; ((oenergy_t *)me->layerEnergy)[l - 1].UpToDate = 0; }
  md->Energy.UpToDate = 0;
  md->Time.Motion += wtime();
}
void EmDee_displace(tEmDee *md, double lambda, double alpha, double dt) { displace_impl(md, lambda, alpha, dt); }
void EmDee_move(tEmDee *md, double lambda, double alpha, double dt) { displace_impl(md, lambda, alpha, dt); }

/* Copy of a body used to look ahead one rotation: EmDeeCode.f90:1150-1163 */
static void virtual_rotation(const obody_t *b, double dt, int mode, double *q) {
  obody_t c = *b;
  c.delta = xcalloc(3 * (size_t)b->NP, sizeof(double));
  double C[12], t[3] = { dt * b->tau[0], dt * b->tau[1], dt * b->tau[2] }, c4[4], in[4];
  matC(c.q, C);
  mul43(C, t, c4);
  for (int k = 0; k < 4; k++) in[k] = c.pi[k] + c4[k];
  body_assign_momenta(&c, in, 4);
  if (mode != 0) rotate_no_squish(&c, dt, mode); else rotate_exact(&c, dt);
  memcpy(q, c.q, 4 * sizeof(double));
  free(c.delta);
}

/* EmDeeCode.f90:1107-1211 */
void EmDee_verlet_step(tEmDee *md, double dt) {
  odata_t *me = ME(*md);
  double Us = 0, Ks_t = 0, Ks_r = 0, dt_2 = HALF * dt;
  int compute = md->Options.Compute, mode = md->Options.RotationMode;
  int nb = me->nbodies, nf = me->nfree;
  double *r0 = xcalloc(3 * (size_t)nb + 1, sizeof(double)), *q0 = xcalloc(4 * (size_t)nb + 1, sizeof(double));
  double *s0 = xcalloc(3 * (size_t)nf + 1, sizeof(double));
  if (compute) {
    for (int i = 1; i <= nb; i++) {
      obody_t *b = &me->body[i];
      double qv[4];
      for (int x = 0; x < 3; x++) r0[3 * (i - 1) + x] = 2.5 * b->rcm[x] + dt_2 * b->invMass * (b->pcm[x] - dt_2 * b->F[x]);
      virtual_rotation(b, -dt, mode, qv);
      for (int k = 0; k < 4; k++) q0[4 * (i - 1) + k] = HALF * qv[k] - 3.0 * b->q[k];
    }
    for (int i = 1; i <= nf; i++) {
      int j = me->free_[i];
      for (int x = 0; x < 3; x++)
        s0[3 * (i - 1) + x] = 2.5 * me->R[3 * (j - 1) + x] + dt_2 * me->invMass[j] * (me->P[3 * (j - 1) + x] - dt_2 * me->F[3 * (j - 1) + x]);
    }
  }
  boost_all(me, ONE, dt_2, me->F, 1, 1);
  move_all(me, ONE, dt, dt, 1, 1, mode);
  EmDee_compute_forces(md);
  boost_all(me, ONE, dt_2, me->F, 1, 1);
  if (compute) {
    double twoKEt[3], twoKEr[3];
    kinetic_energies(me, 1, 1, twoKEt, twoKEr);
    for (int i = 1; i <= nb; i++) {
      obody_t *b = &me->body[i];
      double rdot[3], qdot[4], qv[4];
      for (int x = 0; x < 3; x++) rdot[x] = 2.5 * b->rcm[x] + dt * b->invMass * (b->pcm[x] + dt_2 * b->F[x]) - r0[3 * (i - 1) + x];
      Ks_t += rdot[0] * b->pcm[0] + rdot[1] * b->pcm[1] + rdot[2] * b->pcm[2];
      virtual_rotation(b, dt, mode, qv);
      for (int k = 0; k < 4; k++) qdot[k] = q0[4 * (i - 1) + k] + 1.5 * b->q[k] + qv[k];
      double qq = qdot[0] * b->q[0] + qdot[1] * b->q[1] + qdot[2] * b->q[2] + qdot[3] * b->q[3], s = 0;
      for (int k = 0; k < 4; k++) s += (qdot[k] - qq * b->q[k]) * b->pi[k];
      Ks_r += s;
      /* tau_b = Bt(q) (C(q) tau) */
      double B[12], C[12], c4[4], tb[3];
      matB(b->q, B); matC(b->q, C);
      mul43(C, b->tau, c4);
      mul34t(B, c4, tb);
      Us += b->invMass * (b->F[0] * b->F[0] + b->F[1] * b->F[1] + b->F[2] * b->F[2]) +
            (b->invMoI[0] * (tb[0] * tb[0]) + b->invMoI[1] * (tb[1] * tb[1]) + b->invMoI[2] * (tb[2] * tb[2]));
    }
    for (int i = 1; i <= nf; i++) {
      int j = me->free_[i];
      double rdot[3];
      for (int x = 0; x < 3; x++)
        rdot[x] = 2.5 * me->R[3 * (j - 1) + x] + dt * me->invMass[j] * (me->P[3 * (j - 1) + x] + dt_2 * me->F[3 * (j - 1) + x]) - s0[3 * (i - 1) + x];
      Ks_t += rdot[0] * me->P[3 * (j - 1)] + rdot[1] * me->P[3 * (j - 1) + 1] + rdot[2] * me->P[3 * (j - 1) + 2];
      const double *Fj = me->F + 3 * (j - 1);
      Us += me->invMass[j] * (Fj[0] * Fj[0] + Fj[1] * Fj[1] + Fj[2] * Fj[2]);
    }
    for (int x = 0; x < 3; x++) { md->Kinetic.TransPart[x] = HALF * twoKEt[x]; md->Kinetic.RotPart[x] = HALF * twoKEr[x]; }
    md->Kinetic.Rotational = md->Kinetic.RotPart[0] + md->Kinetic.RotPart[1] + md->Kinetic.RotPart[2];
    md->Kinetic.Kinetic = (md->Kinetic.TransPart[0] + md->Kinetic.TransPart[1] + md->Kinetic.TransPart[2]) + md->Kinetic.Rotational;
    md->Energy.ShadowPotential = md->Energy.ShadowPotential - dt * dt * Us / 24.0;
    md->Kinetic.ShadowRotational = Ks_r / (6.0 * dt);
    md->Kinetic.ShadowKinetic = (Ks_t + Ks_r) / (6.0 * dt);
  }
  md->Kinetic.UpToDate = compute;
  free(r0); free(q0); free(s0);
}

/* EmDeeCode.f90:1281-1395 */
void EmDee_rdf(tEmDee md, int bins, double Rc, int pairs, int itype[], int jtype[], double g[]) {
  const char *task = "radial distribution calculation";
  const double Pi4_3 = 4.188790204786391;
  odata_t *me = ME(md);
  int nt = me->ntypes;
  for (int p = 0; p < pairs; p++) {
    int ij[2] = { itype[p], jtype[p] };
    if (!ranged(ij, 2, nt)) oerror(task, "at least one provided type index is out of range");
  }
  int *hasPair = xcalloc(nt + 1, sizeof(int)), *pairOn = xcalloc((nt + 1) * (nt + 1), sizeof(int)), maxtype = 0;
  for (int p = 0; p < pairs; p++) {
    hasPair[itype[p]] = hasPair[jtype[p]] = 1;
    pairOn[itype[p] * (nt + 1) + jtype[p]] = pairOn[jtype[p] * (nt + 1) + itype[p]] = 1;
    if (itype[p] > maxtype) maxtype = itype[p];
    if (jtype[p] > maxtype) maxtype = jtype[p];
  }
  int npt = symm1D(maxtype, maxtype);
  long *count = xcalloc((size_t)bins * npt, sizeof(long));
  double invL = ONE / *me->Lbox, invL2 = invL * invL, Rc2 = Rc * Rc * invL2, binsByRc = bins / (Rc * invL);
  int N = me->natoms;
  double *Rs = xcalloc(3 * (size_t)N, sizeof(double));
  for (int k = 0; k < 3 * N; k++) Rs[k] = invL * me->R[k];
  for (int t = 1; t <= me->nthreads; t++) {
    olist_t *nb = &me->neighbor[t - 1];
    const int *last = (Rc < 1.0001 * me->InRc) ? nb->middle : nb->last;
    int fa = me->cellFirst[me->threadCellFirst[t]], la = me->cellLast[me->threadCellLast[t]];
    for (int k = fa; k <= la; k++) {
      int i = me->cellItem[k], it = me->atomType[i];
      if (!hasPair[it]) continue;
      for (int m = nb->first[i]; m <= last[i]; m++) {
        int j = nb->item[m], jt = me->atomType[j];
        if (!pairOn[it * (nt + 1) + jt]) continue;
        double r2 = 0;
        for (int x = 0; x < 3; x++) { double d = Rs[3 * (i - 1) + x] - Rs[3 * (j - 1) + x]; d = d - anint(d); r2 += d * d; }
        if (r2 < Rc2) {
          int bin = (int)(sqrt(r2) * binsByRc);
          count[(size_t)bin + (size_t)bins * (symm1D(it, jt) - 1)]++;
        }
      }
    }
  }
  double norm = Pi4_3 * pow(Rc * invL / bins, 3);
  for (int p = 0; p < pairs; p++) {
    int i = itype[p], j = jtype[p];
    long Ni = 0, Nj = 0;
    for (int a = 1; a <= N; a++) { Ni += me->atomType[a] == i; Nj += me->atomType[a] == j; }
    for (int b = 0; b < bins; b++) {
      double r = count[(size_t)b + (size_t)bins * (symm1D(i, j) - 1)] / norm / (3 * (b + 1) * b + 1);
      g[(size_t)b + (size_t)bins * p] = (i == j ? TWO * r : r) / ((double)Ni * (double)Nj);
    }
  }
  free(hasPair); free(pairOn); free(count); free(Rs);
}
