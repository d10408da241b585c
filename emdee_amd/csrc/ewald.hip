// ewald.hip -- reciprocal-space Ewald sum on the GPU (kspace_ewald.f90:188-314).
//
// The reference stores qeikr(k, j) = q_j e^{i k.r_j} for every k-vector and
// charge (nvecs x Ncharged complex FP64: 1.4 GB for 12,288 charges and 7,073
// vectors) and reads it twice.  Here nothing of that size exists:
//   * k_ew_tables: per charge, the powers e^{2 pi i p s_x}, p = 0..nmax, for
//     each axis by the reference's recursion (kspace_ewald.f90:232-246):
//     3 (nmax+1) complex numbers per charge, SoA over charges.
//   * the vectors come in rows along n3 (kspace.h).  A thread owns a charge;
//     at a row start it forms q G1(n1) G2(n2) G3(n3) from the tables, inside
//     a row it multiplies by e^{2 pi i s_z}.  The k data are the same for all
//     lanes of a wavefront, so they are scalar (SGPR) loads.
//   * k_ew_structure: S_I(k) = sum over charges of type I; each wavefront
//     sums 4 x 64 charges of one type per k (registers, then a shuffle tree)
//     and writes one partial per k; k_ew_sigma adds the partials in a fixed
//     order (deterministic), forms sigma = S lambda and the energy per k.
//   * k_ew_forces: F_j = sum_k vec_k (sigma_r(k) Im q_j e^{ik.r_j} -
//     Re q_j e^{ik.r_j} sigma_i(k)) (kspace_ewald.f90:297), with the k range
//     split over blocks for occupancy; k_ew_force_add adds the partial
//     forces in a fixed order into F (cell-sorted slots).
#include <hip/hip_runtime.h>

#include <cmath>

#include "device_state.h"
#include "kspace.h"

namespace emdee {

namespace {
const double TWO_PI = 2.0 * 3.14159265358979324;   // 2 pi with global.f90:38
constexpr int EW_BLOCK = 64;                        // one wavefront per block
constexpr int EW_ATOMS = EWALD_GROUP / EW_BLOCK;    // charges per thread in k_ew_structure
constexpr int EW_TARGET_WAVES = 16384;              // k-range split target of the structure factors
constexpr int EW_TARGET_WAVES_F = 16384;            // ... and of the forces
constexpr int EW_ADD_LANES = 16;                    // lanes per charge adding the partial forces
}  // namespace

struct EwaldState {
  int nm = 0, nvecs = 0, nt = 0, nl = 0, nslots = 0, ngroups = 0;
  int ksS = 1, kchS = 0, ksF = 1, kchF = 0;
  DBuf<int4> kn;
  DBuf<double4> kv;
  DBuf<int> item, groupType, typeGroups;
  DBuf<double> q, lambda, Fpart, part, selfc;
  DBuf<double> vfac;      // per k: 1 - k^2 / (2 alpha^2), when the virial of the sum is wanted
  DBuf<double2> T, Spart, S, sigma;
  std::vector<int> hostTypeGroups;
};

static EwaldState* ewald_of(Device* d) {
  if (!d->ewald) d->ewald = new EwaldState();
  return d->ewald;
}

void dev_ewald_release(Device* d) {
  EwaldState* e = d->ewald;
  if (!e) return;
  e->kn.release(); e->kv.release(); e->item.release(); e->groupType.release(); e->typeGroups.release();
  e->q.release(); e->lambda.release(); e->Fpart.release(); e->part.release(); e->selfc.release();
  e->vfac.release(); e->T.release(); e->Spart.release(); e->S.release(); e->sigma.release();
  delete e;
  d->ewald = nullptr;
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// e^{2 pi i p s} for p = 0..nm on each axis, kspace_ewald.f90:209-246.
// T[(x*(nm+1) + p)*nslots + slot].
__global__ void k_ew_tables(double2* __restrict__ T, const int* __restrict__ item, const int* __restrict__ iperm,
                            const double* __restrict__ R, double L, int nm, int nslots, int Npad) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nslots) return;
  int i = item[j];
  int s = i >= 0 ? iperm[i] : -1;
  for (int x = 0; x < 3; x++) {
    double2 Z = make_double2(1.0, 0.0);
    if (s >= 0) {
      double th = TWO_PI * (R[x * Npad + s] / L);
      double sn, cs;
      sincos(th, &sn, &cs);
      Z = make_double2(cs, sn);
    }
    double2* Tx = T + (size_t)x * (nm + 1) * nslots + j;
    Tx[0] = make_double2(1.0, 0.0);
    double2 b = Z;
    for (int p = 1; p <= nm; p++) {
      Tx[(size_t)p * nslots] = b;
      b = cmul(b, Z);
    }
  }
}

// q G1(n1) G2(n2) G3(n3), with G(-p) = conj(G(p)); kspace_ewald.f90:216-219.
__device__ __forceinline__ double2 row_start(const double2* __restrict__ T, double q, int4 n, int j, int nm,
                                             int nslots) {
  size_t stride = (size_t)(nm + 1) * nslots;
  double2 g1 = T[(size_t)n.x * nslots + j];
  double2 g2 = T[stride + (size_t)abs(n.y) * nslots + j];
  double2 g3 = T[2 * stride + (size_t)abs(n.z) * nslots + j];
  if (n.y < 0) g2.y = -g2.y;
  if (n.z < 0) g3.y = -g3.y;
  double2 v = make_double2(q * g1.x, q * g1.y);
  v = cmul(v, g2);
  return cmul(v, g3);
}

// Partial structure factors: Spart[g*nvecs + k] = sum over the EWALD_GROUP
// slots of group g of q e^{ik.r}.  Grid (ngroups, ksS), one wavefront each.
// Each lane sums its EW_ATOMS charges for EW_KB consecutive k-vectors in
// registers; the 64 lane sums of those k are then added through LDS, one
// lane per (k, real/imaginary part) in lane order, instead of a shuffle
// tree per k.
constexpr int EW_KB = 16;
__global__ void __launch_bounds__(EW_BLOCK) k_ew_structure(double2* __restrict__ Spart, const double2* __restrict__ T,
                                                           const double* __restrict__ q, const int4* __restrict__ kn,
                                                           int nm, int nslots, int nvecs, int kch) {
  __shared__ double red[2 * EW_KB][EW_BLOCK + 1];
  int g = blockIdx.x, lane = threadIdx.x;
  int k0 = blockIdx.y * kch, k1 = min(nvecs, k0 + kch);
  int j[EW_ATOMS];
  double qa[EW_ATOMS];
  double2 z[EW_ATOMS], v[EW_ATOMS];
  size_t zoff = (2 * (size_t)(nm + 1) + 1) * nslots;   // G3(1) = e^{2 pi i s_z}
#pragma unroll
  for (int a = 0; a < EW_ATOMS; a++) {
    j[a] = g * EWALD_GROUP + a * EW_BLOCK + lane;
    qa[a] = q[j[a]];
    z[a] = T[zoff + j[a]];
    v[a] = make_double2(0.0, 0.0);
  }
  double* out = reinterpret_cast<double*>(Spart + (size_t)g * nvecs);
  for (int kb = k0; kb < k1; kb += EW_KB) {
#pragma unroll
    for (int t = 0; t < EW_KB; t++) {
      int k = kb + t;
      double sr = 0.0, si = 0.0;
      if (k < k1) {
        int4 n = kn[k];
        if (k == k0 || n.w) {
#pragma unroll
          for (int a = 0; a < EW_ATOMS; a++) v[a] = row_start(T, qa[a], n, j[a], nm, nslots);
        } else {
#pragma unroll
          for (int a = 0; a < EW_ATOMS; a++) v[a] = cmul(v[a], z[a]);
        }
#pragma unroll
        for (int a = 0; a < EW_ATOMS; a++) { sr += v[a].x; si += v[a].y; }
      }
      red[2 * t][lane] = sr;
      red[2 * t + 1][lane] = si;
    }
    __syncthreads();
    if (lane < 2 * EW_KB && kb + (lane >> 1) < k1) {
      double s0 = 0.0, s1 = 0.0;
#pragma unroll 8
      for (int l = 0; l < EW_BLOCK; l += 2) { s0 += red[lane][l]; s1 += red[lane][l + 1]; }
      out[2 * (size_t)kb + lane] = s0 + s1;
    }
    __syncthreads();
  }
}

// S_I(k) (kspace_ewald.f90:226-228), sigma(k) = S(k) lambda (:284) and the
// energy prefac * sum_I S_I . sigma_I (:287), one wavefront per k: lane g
// adds the partials of groups g, g + 64, ... of each charged type, a fixed
// shuffle tree completes S_I(k), and lane 0 forms sigma and the energy.  The
// four k of a block are added in a fixed order into one partial per block.
// With vfac (a partial k set standing in for part of the sum), the partials
// also carry the virial of the energy, sum_k E_k (1 - k^2/(2 alpha^2)) =
// -dE/dln(lambda) when positions and box scale by lambda: part[2 b + 1].
constexpr int EW_SIGMA_MAXT = 8;
__global__ void __launch_bounds__(BLOCK) k_ew_sigma(double2* __restrict__ S, double2* __restrict__ sigma,
                                                    double* __restrict__ part, const double2* __restrict__ Spart,
                                                    const int* __restrict__ typeGroups, const double* __restrict__ lam,
                                                    const double4* __restrict__ kv, const double* __restrict__ vfac,
                                                    int nt, int nvecs) {
  __shared__ double red[2][BLOCK / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int k = blockIdx.x * (BLOCK / 64) + wv;
  double e = 0.0, v = 0.0;
  if (k < nvecs) {
    double sr[EW_SIGMA_MAXT], si[EW_SIGMA_MAXT];
    for (int t = 0; t < nt; t++) {
      double a = 0.0, b = 0.0;
      for (int g = typeGroups[t] + lane; g < typeGroups[t + 1]; g += 64) {
        double2 p = Spart[(size_t)g * nvecs + k];
        a += p.x;
        b += p.y;
      }
      for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
      }
      sr[t] = a;
      si[t] = b;
    }
    if (lane == 0) {
      double s = 0.0;
      for (int t = 0; t < nt; t++) S[(size_t)t * nvecs + k] = make_double2(sr[t], si[t]);
      for (int t2 = 0; t2 < nt; t2++) {
        double gr = 0.0, gi = 0.0;
        for (int t1 = 0; t1 < nt; t1++) {
          gr += sr[t1] * lam[t1 + nt * t2];
          gi += si[t1] * lam[t1 + nt * t2];
        }
        sigma[(size_t)t2 * nvecs + k] = make_double2(gr, gi);
        s += sr[t2] * gr + si[t2] * gi;
      }
      e = kv[k].w * s;
      if (vfac) v = e * vfac[k];
    }
  }
  if (lane == 0) { red[0][wv] = e; red[1][wv] = v; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double se = 0.0, sv = 0.0;
    for (int w = 0; w < BLOCK / 64; w++) { se += red[0][w]; sv += red[1][w]; }
    if (vfac) { part[2 * blockIdx.x] = se; part[2 * blockIdx.x + 1] = sv; }
    else part[blockIdx.x] = se;
  }
}

// Partial k-space forces over the k range of blockIdx.y:
// Fpart[(y*4 + x)*nslots + j], x = 0..2; x = 3 holds the atom's share of the
// reciprocal energy, sum_k prefac Re(q_j e^{ik.r_j} conj(sigma_I(k))), whose
// sum over atoms is kspace_ewald.f90:287.  Every wavefront's slots hold one
// charged type.
__global__ void __launch_bounds__(EW_BLOCK) k_ew_forces(double* __restrict__ Fpart, const double2* __restrict__ T,
                                                        const double* __restrict__ q, const int4* __restrict__ kn,
                                                        const double4* __restrict__ kv,
                                                        const double2* __restrict__ sigma,
                                                        const int* __restrict__ groupType, int nm, int nslots,
                                                        int nvecs, int kch) {
  int j = blockIdx.x * EW_BLOCK + threadIdx.x;
  int k0 = blockIdx.y * kch, k1 = min(nvecs, k0 + kch);
  int t = groupType[(blockIdx.x * EW_BLOCK) / EWALD_GROUP];
  const double2* sg = sigma + (size_t)t * nvecs;
  double qj = q[j];
  size_t zoff = (2 * (size_t)(nm + 1) + 1) * nslots;   // G3(1) = e^{2 pi i s_z}
  double2 z = T[zoff + j], va = make_double2(0.0, 0.0), vb = va;
  double fx = 0.0, fy = 0.0, fz = 0.0, er = 0.0;
  // two halves of the k range advance together: two independent recursions
  // in flight per lane
  const int kh = k0 + (k1 - k0 + 1) / 2;
  auto term = [&](int k, const double2& v) {
    double2 s = sg[k];
    double4 w = kv[k];
    double c = s.x * v.y - v.x * s.y;
    fx += w.x * c;
    fy += w.y * c;
    fz += w.z * c;
    er += w.w * (s.x * v.x + s.y * v.y);
  };
#pragma unroll 4
  for (int t = 0; k0 + t < kh; t++) {
    int ka = k0 + t, kb = kh + t;
    int4 na = kn[ka];
    va = (t == 0 || na.w) ? row_start(T, qj, na, j, nm, nslots) : cmul(va, z);
    if (kb < k1) {
      int4 nb = kn[kb];
      vb = (t == 0 || nb.w) ? row_start(T, qj, nb, j, nm, nslots) : cmul(vb, z);
    }
    term(ka, va);
    if (kb < k1) term(kb, vb);
  }
  size_t o = (size_t)blockIdx.y * 4 * nslots + j;
  Fpart[o] = fx;
  Fpart[o + nslots] = fy;
  Fpart[o + 2 * (size_t)nslots] = fz;
  Fpart[o + 3 * (size_t)nslots] = er;
}

// F(slot of atom) += sum over k ranges (fixed order), kspace_ewald.f90:297;
// with eatom, also the atom's reciprocal and self energies
// (self: -lambda_II alpha q^2 / sqrt(pi), modelClass_kspace.f90:223-229).
// EW_ADD_LANES lanes per charge: lane u adds the ranges u, u + 16, ... in
// order, and a fixed butterfly adds the lanes' sums.
__global__ void k_ew_force_add(double* __restrict__ F, double* __restrict__ eatom, const double* __restrict__ Fpart,
                               const int* __restrict__ item, const int* __restrict__ iperm,
                               const double* __restrict__ q, const int* __restrict__ groupType,
                               const double* __restrict__ selfc, int nslots, int ks, int Npad) {
  int gt = blockIdx.x * blockDim.x + threadIdx.x;
  int j = gt / EW_ADD_LANES, u = gt % EW_ADD_LANES;
  bool ok = j < nslots;
  double f[4] = { 0.0, 0.0, 0.0, 0.0 };
  if (ok) {
    for (int y = u; y < ks; y += EW_ADD_LANES)
#pragma unroll
      for (int x = 0; x < 4; x++) f[x] += Fpart[((size_t)y * 4 + x) * nslots + j];
  }
#pragma unroll
  for (int off = EW_ADD_LANES / 2; off > 0; off >>= 1)
#pragma unroll
    for (int x = 0; x < 4; x++) f[x] += __shfl_xor(f[x], off, EW_ADD_LANES);
  if (!ok || u != 0) return;
  int i = item[j];
  if (i < 0) return;
  int s = iperm[i];
  for (int x = 0; x < 3; x++) F[x * (size_t)Npad + s] += f[x];
  if (eatom) eatom[s] += f[3] + selfc[groupType[j / EWALD_GROUP]] * q[j] * q[j];
}

static void ewald_state_setup(EwaldState* e, const EwaldSetup& s) {
  e->nm = s.nm; e->nvecs = s.nvecs; e->nt = s.nt; e->nl = s.nl; e->nslots = s.nslots;
  e->ngroups = s.nslots / EWALD_GROUP;
  e->hostTypeGroups = s.typeGroups;
  auto put = [&](auto& buf, const auto& host) {
    buf.alloc(host.size());
    HIP_CHECK(hipMemcpy(buf.p, host.data(), host.size() * sizeof(host[0]), hipMemcpyHostToDevice));
  };
  for (int k = 0; k < s.nvecs; k++)
    if (s.kn[4 * k] < 0 || s.kn[4 * k] > s.nm || std::abs(s.kn[4 * k + 1]) > s.nm || std::abs(s.kn[4 * k + 2]) > s.nm)
      fatal("kspace set-up", "k-vector outside the half space of the table");
  if (s.nslots % EWALD_GROUP || (int)s.groupType.size() != s.nslots / EWALD_GROUP)
    fatal("kspace set-up", "charge slots are not grouped");
  std::vector<int4> kn(s.nvecs);
  std::vector<double4> kv(s.nvecs);
  for (int k = 0; k < s.nvecs; k++) {
    kn[k] = make_int4(s.kn[4 * k], s.kn[4 * k + 1], s.kn[4 * k + 2], s.kn[4 * k + 3]);
    kv[k] = make_double4(s.kv[4 * k], s.kv[4 * k + 1], s.kv[4 * k + 2], s.kv[4 * k + 3]);
  }
  put(e->kn, kn);
  put(e->kv, kv);
  put(e->item, s.slotItem);
  put(e->q, s.slotQ);
  put(e->groupType, s.groupType);
  put(e->typeGroups, s.typeGroups);
  put(e->lambda, s.lambda);
  // k-range splits: enough wavefronts to fill the chip
  e->ksS = std::max(1, std::min(s.nvecs, EW_TARGET_WAVES / std::max(1, e->ngroups)));
  e->kchS = div_up(s.nvecs, e->ksS);
  e->ksS = div_up(s.nvecs, e->kchS);
  int nwF = s.nslots / EW_BLOCK;
  e->ksF = std::max(1, std::min(s.nvecs, EW_TARGET_WAVES_F / std::max(1, nwF)));
  e->kchF = div_up(s.nvecs, e->ksF);
  e->ksF = div_up(s.nvecs, e->kchF);
  e->T.alloc(3 * (size_t)(s.nm + 1) * s.nslots);
  e->Spart.alloc((size_t)e->ngroups * s.nvecs);
  e->S.alloc((size_t)s.nt * s.nvecs);
  e->sigma.alloc((size_t)s.nt * s.nvecs);
  e->Fpart.alloc((size_t)e->ksF * 4 * s.nslots);
  if (s.nt > EW_SIGMA_MAXT) fatal("kspace set-up", "more than 8 charged types");
  e->part.alloc(2 * (size_t)div_up(s.nvecs, BLOCK / 64));
}

void dev_ewald_setup(Device* d, const EwaldSetup& s) {
  EwaldState* e = ewald_of(d);
  ewald_state_setup(e, s);
  std::vector<double> selfc((size_t)s.nl * s.nt);
  for (int l = 0; l < s.nl; l++)
    for (int t = 0; t < s.nt; t++)
      selfc[(size_t)l * s.nt + t] = -s.alpha * s.lambda[(size_t)l * s.nt * s.nt + t * s.nt + t] / std::sqrt(3.14159265358979324);
  e->selfc.alloc(selfc.size());
  HIP_CHECK(hipMemcpy(e->selfc.p, selfc.data(), selfc.size() * sizeof(double), hipMemcpyHostToDevice));
}

// One evaluation of the sum over the state's k-vectors with the type-pair
// weights lam and self-energy coefficients selfc (null: none): forces added
// to F (slot order), atom shares to eatom (may be null); returns the energy
// and, with the state's virial factors, its virial in *W.
static unsigned long long ewald_launch(Device* d, EwaldState* e, double* F, const double* lam, const double* selfc,
                                       double* eatom, double L, bool virial, int slot) {
  // shapes the kernels assume
  if (e->nslots % EWALD_GROUP || e->ngroups * EWALD_GROUP != e->nslots || e->kchS * e->ksS < e->nvecs ||
      e->kchF * e->ksF < e->nvecs || e->nm < 1)
    fatal("kspace computation", "inconsistent k-space set-up");
  if (eatom && !selfc) fatal("kspace computation", "atom energies need the self-energy coefficients");
  int Np = d->Npad;
  const double* vf = virial ? e->vfac.p : nullptr;
  if (virial && !vf) fatal("kspace computation", "no virial factors for this k set");
  hipLaunchKernelGGL(k_ew_tables, dim3(div_up(e->nslots, BLOCK)), dim3(BLOCK), 0, d->stream, e->T.p, e->item.p,
                     d->iperm.p, d->R.p, L, e->nm, e->nslots, Np);
  launch_check("ewald tables");
  hipLaunchKernelGGL(k_ew_structure, dim3(e->ngroups, e->ksS), dim3(EW_BLOCK), 0, d->stream, e->Spart.p, e->T.p,
                     e->q.p, e->kn.p, e->nm, e->nslots, e->nvecs, e->kchS);
  launch_check("ewald structure factors");
  int gk = div_up(e->nvecs, BLOCK / 64);
  hipLaunchKernelGGL(k_ew_sigma, dim3(gk), dim3(BLOCK), 0, d->stream, e->S.p, e->sigma.p, e->part.p, e->Spart.p,
                     e->typeGroups.p, lam, e->kv.p, vf, e->nt, e->nvecs);
  launch_check("ewald sigma");
  // energy (and virial) straight into pinned host memory: the host reads
  // them as soon as they arrive, while the force kernels below still run
  const unsigned long long seq = ++d->seq;
  if (vf)
    hipLaunchKernelGGL(k_reduce_post<2>, dim3(1), dim3(BLOCK), 0, d->stream, d->hpostDev + slot, d->hseqDev, seq,
                       e->part.p, gk);
  else
    hipLaunchKernelGGL(k_reduce_post<1>, dim3(1), dim3(BLOCK), 0, d->stream, d->hpostDev + slot, d->hseqDev, seq,
                       e->part.p, gk);
  hipLaunchKernelGGL(k_ew_forces, dim3(e->nslots / EW_BLOCK, e->ksF), dim3(EW_BLOCK), 0, d->stream, e->Fpart.p,
                     e->T.p, e->q.p, e->kn.p, e->kv.p, e->sigma.p, e->groupType.p, e->nm, e->nslots, e->nvecs,
                     e->kchF);
  launch_check("ewald forces");
  hipLaunchKernelGGL(k_ew_force_add, dim3(div_up((long)e->nslots * EW_ADD_LANES, BLOCK)), dim3(BLOCK), 0, d->stream, F,
                     eatom, e->Fpart.p,
                     e->item.p, d->iperm.p, e->q.p, e->groupType.p, selfc, e->nslots, e->ksF, Np);
  launch_check("ewald force add");
  return seq;
}

// the same, read once its energy (and virial) have been posted
static double ewald_eval(Device* d, EwaldState* e, double* F, const double* lam, const double* selfc, double* eatom,
                         double L, double* W, int slot) {
  dev_post_wait(d, ewald_launch(d, e, F, lam, selfc, eatom, L, W != nullptr, slot));
  if (W) *W = d->hpost[slot + 1];
  return d->hpost[slot];
}

bool dev_ewald_launch(Device* d, int layer, double L) {
  EwaldState* e = d->ewald;
  if (!e || e->nvecs == 0 || e->nslots == 0) return false;
  double* F = d->F.p + 3 * (size_t)d->Npad * layer;
  ewald_launch(d, e, F, e->lambda.p + (size_t)layer * e->nt * e->nt, e->selfc.p + (size_t)layer * e->nt,
               d->wantAtomE ? d->eatom.p : nullptr, L, false, POST_EWALD);
  return true;
}

double dev_ewald_energy(const Device* d) { return d->hpost[POST_EWALD]; }

double dev_ewald_compute(Device* d, int layer, double L) {
  if (!dev_ewald_launch(d, layer, L)) return 0.0;
  dev_wait_posts(d);
  return dev_ewald_energy(d);
}

// A reciprocal-space sum over a partial k set (set-up by lowk_setup): one
// weight table (layer 0 of s.lambda), no self energy, with its virial.
EwaldState* dev_lowk_create(const EwaldSetup& s) {
  if (s.nl != 1) fatal("kspace set-up", "a partial k set has one weight table");
  EwaldState* e = new EwaldState();
  ewald_state_setup(e, s);
  std::vector<double> vf(s.nvecs), zero(s.nt, 0.0);
  for (int k = 0; k < s.nvecs; k++) {
    // kv = (2 prefac k, prefac): k^2 = |kv.xyz|^2 / (2 prefac)^2
    double p2 = 2.0 * s.kv[4 * k + 3];
    double kx = s.kv[4 * k] / p2, ky = s.kv[4 * k + 1] / p2, kz = s.kv[4 * k + 2] / p2;
    vf[k] = 1.0 - (kx * kx + ky * ky + kz * kz) / (2.0 * s.alpha * s.alpha);
  }
  e->vfac.alloc(vf.size());
  HIP_CHECK(hipMemcpy(e->vfac.p, vf.data(), vf.size() * sizeof(double), hipMemcpyHostToDevice));
  e->selfc.alloc(s.nt);
  HIP_CHECK(hipMemcpy(e->selfc.p, zero.data(), zero.size() * sizeof(double), hipMemcpyHostToDevice));
  return e;
}

double dev_lowk_compute(Device* d, EwaldState* e, double* F, double* eatom, double L, double* W) {
  if (!e || e->nvecs == 0 || e->nslots == 0) { *W = 0.0; return 0.0; }
  return ewald_eval(d, e, F, e->lambda.p, e->selfc.p, eatom, L, W, POST_LOWK);
}

bool dev_lowk_launch(Device* d, EwaldState* e, double* F, double* eatom, double L) {
  if (!e || e->nvecs == 0 || e->nslots == 0) return false;
  ewald_launch(d, e, F, e->lambda.p, e->selfc.p, eatom, L, true, POST_LOWK);
  return true;
}

double dev_lowk_result(const Device* d, double* W) {
  *W = d->hpost[POST_LOWK + 1];
  return d->hpost[POST_LOWK];
}

int dev_lowk_vectors(const EwaldState* e) { return e ? e->nvecs : 0; }

void dev_lowk_release(EwaldState* e) {
  if (!e) return;
  e->kn.release(); e->kv.release(); e->item.release(); e->groupType.release(); e->typeGroups.release();
  e->q.release(); e->lambda.release(); e->Fpart.release(); e->part.release(); e->selfc.release();
  e->vfac.release(); e->T.release(); e->Spart.release(); e->S.release(); e->sigma.release();
  delete e;
}

}  // namespace emdee
