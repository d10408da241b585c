// pair_eval.h -- device-side evaluation of one non-bonded pair.
//
// Restates, per pair, the body of the reference pair loop
// (src/compute.f90:46-92) with the model bodies of pair_lj_cut.f90:73-86,
// pair_softcore_cut.f90:86-101, coul_*.f90 and the modifier block of
// src/apply_modifier.f90:1-60.  The reference dispatches with `select type`
// per pair; here the record kinds are runtime values that are uniform across a
// wavefront for single-model systems, and the hot single-model cases are
// compiled as template specialisations in device.hip.
//
// T is the arithmetic type of the model evaluation: double for the plain
// form, float for the mixed form (the caller accumulates in double).
#pragma once
#include <hip/hip_runtime.h>
#include "engine.h"

namespace emdee {

template <typename T>
__device__ __forceinline__ T rcp(T x) { return T(1) / x; }
template <>
__device__ __forceinline__ float rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }

// a / b and sqrt: IEEE for double (the plain form keeps the reference
// arithmetic); the hardware reciprocal and square root (about 1 ulp) for the
// float terms of the mixed form
template <typename T>
__device__ __forceinline__ T quo(T a, T b) { return a / b; }
template <>
__device__ __forceinline__ float quo<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float sqrt_fast(float x) { return __builtin_amdgcn_sqrtf(x); }

template <typename T>
__device__ __forceinline__ T uerfc_dev(T x, T expmx2) {
  // Abramowitz & Stegun 7.1.26 (math.f90:35-40, 685-691)
  const T a1 = T(0.254829592), a2 = T(-0.284496736), a3 = T(1.421413741),
          a4 = T(-1.453152027), a5 = T(1.061405429), p = T(0.327591100);
  T t = rcp<T>(T(1) + p * x);
  return t * (a1 + t * (a2 + t * (a3 + t * (a4 + t * a5)))) * expmx2;
}

template <typename T> __device__ __forceinline__ T dexp(T x) { return exp(x); }
template <> __device__ __forceinline__ float dexp<float>(float x) { return __expf(x); }

// Pair model E and W (W = -r dE/dr).
template <typename T>
__device__ __forceinline__ void pair_model(const DevPair& p, T invR2, T& E, T& W) {
  if (p.kind == PAIR_LJ_CUT) {
    T sr2 = T(p.c2) * invR2, sr6 = sr2 * sr2 * sr2, sr12 = sr6 * sr6;
    E = T(p.c0) * (sr12 - sr6);
    W = T(p.c1) * (sr12 + sr12 - sr6);
  } else if (p.kind == PAIR_SOFTCORE_CUT) {
    T rsig2 = quo<T>(T(p.c2), invR2), rsig6 = rsig2 * rsig2 * rsig2;
    T sinv = rcp<T>(rsig6 + T(p.c3)), sinvSq = sinv * sinv, sinvCb = sinv * sinvSq;
    E = T(p.c0) * (sinvSq - sinv);
    W = T(p.c1) * rsig6 * (sinvCb + sinvCb - sinvSq);
  } else {
    E = T(0);
    W = T(0);
  }
}

// apply_modifier.f90: the W update is the same with and without energies;
// E is only meaningful when energies are requested.
template <typename T>
__device__ __forceinline__ void apply_modifier(int mod, T eshift, T fshift, T factor, T Rm2fac,
                                               T invR, T invR2, T& E, T& W) {
  switch (mod) {
    case MOD_SHIFTED:
      E = E + eshift;
      break;
    case MOD_SHIFTED_FORCE: {
      T rFc = quo<T>(fshift, invR);
      W = W - rFc;
      E = E + eshift + rFc;
      break;
    }
    case MOD_SMOOTHED: case MOD_SHIFTED_SMOOTHED: {
      E = E + eshift;
      T r2fac = quo<T>(factor, invR);
      if (r2fac > Rm2fac) {
        T u = r2fac - Rm2fac, u2 = u * u, u3 = u * u2;
        T G = T(1) + u3 * (T(15) * u - T(6) * u2 - T(10));
        T WG = T(-30) * u2 * (T(2) * u - u2 - T(1)) * r2fac;
        W = W * G + E * WG;
        E = E * G;
      }
      break;
    }
    case MOD_SQUARE_SMOOTHED: case MOD_SHIFTED_SQUARE_SMOOTHED: {
      E = E + eshift;
      T r2fac = quo<T>(factor, invR2);
      if (r2fac > Rm2fac) {
        T u = r2fac - Rm2fac, u2 = u * u, u3 = u * u2;
        T G = T(1) + u3 * (T(15) * u - T(6) * u2 - T(10));
        T WG = T(-60) * u2 * (T(2) * u - u2 - T(1)) * r2fac;
        W = W * G + E * WG;
        E = E * G;
      }
      break;
    }
    default: break;
  }
}

// Coulomb model E and W (without the charge product).  ENERGY selects the
// reference's "compute" body; otherwise its "virial" body, which differs for
// coul_square_smoothed (coul_square_smoothed.f90:131-148) and leaves W
// untouched for coul_none (the generated virial dispatch has no coul_none
// case, make_virial_compute.sh:26-30).
template <typename T, bool ENERGY>
__device__ __forceinline__ void coul_model(const DevCoul& c, T invR, T invR2, T& E, T& W) {
  switch (c.kind) {
    case COUL_CUT:
      E = invR; W = invR;
      break;
    case COUL_SF: {
      T rFc = quo<T>(T(c.fshift), invR);
      E = invR + T(c.eshift) + rFc;
      W = invR - rFc;
      break;
    }
    case COUL_DAMPED: case COUL_LONG: {
      T x = quo<T>(T(c.alpha), invR), expmx2 = dexp<T>(-x * x);
      E = uerfc_dev<T>(x, expmx2) * invR;
      W = E + T(c.beta) * expmx2;
      break;
    }
    case COUL_DAMPED_SMOOTHED: {
      T r = rcp<T>(invR), x = T(c.alpha) * r, expmx2 = dexp<T>(-x * x);
      E = uerfc_dev<T>(x, expmx2) * invR;
      W = E + T(c.beta) * expmx2;
      if (r > T(c.Rm)) {
        T u = T(c.factor) * (r - T(c.Rm)), u2 = u * u, u3 = u * u2;
        T G = T(1) + u3 * (T(15) * u - T(6) * u2 - T(10));
        T WG = T(-30) * u2 * (T(2) * u - u2 - T(1)) * T(c.factor) * r;
        W = W * G + E * WG;
        E = E * G;
      }
      break;
    }
    case COUL_DAMPED_SQUARE_SMOOTHED: {
      T x = quo<T>(T(c.alpha), invR), expmx2 = dexp<T>(-x * x);
      E = uerfc_dev<T>(x, expmx2) * invR;
      W = E + T(c.beta) * expmx2;
      if (invR < T(c.invRm)) {
        T r2 = rcp<T>(invR2), u = T(c.factor) * (r2 - T(c.Rm2)), u2 = u * u, u3 = u * u2;
        T G = T(1) + u3 * (T(15) * u - T(6) * u2 - T(10));
        T WG = T(-60) * u2 * (T(2) * u - u2 - T(1)) * T(c.factor) * r2;
        W = W * G + E * WG;
        E = E * G;
      }
      break;
    }
    case COUL_SQUARE_SMOOTHED: case COUL_SHIFTED_SQUARE_SMOOTHED: {
      W = invR;
      E = (c.kind == COUL_SQUARE_SMOOTHED) ? W : W + T(c.eshift);
      if (invR < T(c.invRm)) {
        T r2 = rcp<T>(invR2), u = T(c.factor) * (r2 - T(c.Rm2)), u2 = u * u, u3 = u * u2;
        T G = T(1) + u3 * (T(15) * u - T(6) * u2 - T(10));
        T WG = T(-60) * u2 * (T(2) * u - u2 - T(1)) * T(c.factor) * r2;
        if (!ENERGY && c.kind == COUL_SQUARE_SMOOTHED) W = W * (G + WG);
        else W = W * G + E * WG;
        E = E * G;
      }
      break;
    }
    default:
      if (ENERGY) { E = T(0); W = T(0); }
      break;
  }
}

}  // namespace emdee
