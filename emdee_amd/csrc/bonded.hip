// bonded.hip -- harmonic bonds and angles (EmDeeData.f90:443-550,
// bond_harmonic.f90:69-81, angle_harmonic.f90:69-79), harmonic dihedrals
// (EmDeeData.f90:554-640) and, when an Ewald solver is active, the erf
// discount of the excluded 1-2 and 1-3 pairs (modelClass_kspace.f90:321-334).
//
// One thread per bond or angle computes its force contributions into a small
// per-structure array; a gather kernel then adds, for every atom, the
// contributions of the structures it belongs to in a fixed order (CSR of
// incidences built on the host).  No atomics: sums are deterministic.
#include <hip/hip_runtime.h>

#include <cmath>

#include "device_state.h"
#include "kspace.h"
#include "pair_eval.h"

namespace emdee {

struct StructState {
  int nb = 0, na = 0, nd = 0;
  DBuf<DevStruct> bonds, angles;
  DBuf<DevDihedral> dihedrals;
  DBuf<double> bF, aF, dF, part, result;
  DBuf<int> incOff, incItem;   // incidences of each original atom: (index << 4) | role
};

enum { ROLE_BOND_I = 0, ROLE_BOND_J = 1, ROLE_ANGLE_I = 2, ROLE_ANGLE_K = 3, ROLE_ANGLE_J = 4, ROLE_DIH_I = 5,
       ROLE_DIH_J = 6, ROLE_DIH_K = 7, ROLE_DIH_L = 8 };

void dev_structs_release(Device* d) {
  StructState* s = d->structs;
  if (!s) return;
  s->bonds.release(); s->angles.release(); s->dihedrals.release(); s->bF.release(); s->aF.release();
  s->dF.release(); s->part.release();
  s->result.release(); s->incOff.release(); s->incItem.release();
  delete s;
  d->structs = nullptr;
}

__device__ __forceinline__ void ew_discount(double alpha, double beta, double rsq, double QiQj, double& E, double& W) {
  double r = sqrt(rsq), x = alpha * r, expmx2 = exp(-x * x);
  E = -QiQj * (1.0 - uerfc_dev<double>(x, expmx2)) / r;
  W = E + QiQj * beta * expmx2;
}

// EmDeeData.f90:443-487.  out per block: {Ebond, Wbond, Ecoul}.
__global__ void __launch_bounds__(BLOCK) k_bonds(double* __restrict__ bF, double* __restrict__ part,
                                                 const DevStruct* __restrict__ bonds, const double* __restrict__ R,
                                                 const int* __restrict__ iperm, const int* __restrict__ type,
                                                 const double* __restrict__ charge, const DevPair* __restrict__ pt,
                                                 int ntypes, double L, int bonded, int kspace, double alpha,
                                                 double beta, int nb, int Npad) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[3] = { 0.0, 0.0, 0.0 };
  if (b < nb) {
    DevStruct s = bonds[b];
    int si = iperm[s.i], sj = iperm[s.j];
    double invL2 = 1.0 / (L * L), Rij[3];
    for (int x = 0; x < 3; x++) {
      Rij[x] = R[x * (size_t)Npad + si] / L - R[x * (size_t)Npad + sj] / L;
      Rij[x] = Rij[x] - round(Rij[x]);
    }
    double invR2 = invL2 / (Rij[0] * Rij[0] + Rij[1] * Rij[1] + Rij[2] * Rij[2]);
    double W = 0.0;
    if (bonded) {
      double E = 0.0;
      if (s.kind == BOND_HARMONIC) {
        double r = 1.0 / sqrt(invR2), delta = r - s.x0;
        E = s.half_k * delta * delta;
        W = s.minus_k * delta * r;
      }
      acc[0] = E;
      acc[1] = W;
    }
    if (kspace) {
      double QiQj = pt[type[si] * ntypes + type[sj]].kCoul * charge[si] * charge[sj], EL, WL;
      ew_discount(alpha, beta, 1.0 / invR2, QiQj, EL, WL);
      acc[2] = EL;
      W = W + WL;
    }
    for (int x = 0; x < 3; x++) bF[3 * (size_t)b + x] = W * invR2 * L * Rij[x];
  }
  block_sum<3>(acc);
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; k++) part[3 * blockIdx.x + k] = acc[k];
}

// EmDeeData.f90:491-550.  aF per angle: force on i, on k, on j.
// out per block: {Eangle, Wangle, Ecoul}.
__global__ void __launch_bounds__(BLOCK) k_angles(double* __restrict__ aF, double* __restrict__ part,
                                                  const DevStruct* __restrict__ angles, const double* __restrict__ R,
                                                  const int* __restrict__ iperm, const int* __restrict__ type,
                                                  const double* __restrict__ charge, const DevPair* __restrict__ pt,
                                                  int ntypes, double L, int bonded, int kspace, double alpha,
                                                  double beta, int na, int Npad) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[3] = { 0.0, 0.0, 0.0 };
  if (m < na) {
    DevStruct s = angles[m];
    int si = iperm[s.i], sj = iperm[s.j], sk = iperm[s.k];
    double avec[3], bvec[3], Fi[3] = { 0, 0, 0 }, Fk[3] = { 0, 0, 0 }, Fik[3] = { 0, 0, 0 };
    for (int x = 0; x < 3; x++) {
      double ri = R[x * (size_t)Npad + si] / L, rj = R[x * (size_t)Npad + sj] / L, rk = R[x * (size_t)Npad + sk] / L;
      avec[x] = ri - rj;
      bvec[x] = rk - rj;
      avec[x] = L * (avec[x] - round(avec[x]));
      bvec[x] = L * (bvec[x] - round(bvec[x]));
    }
    if (bonded) {
      double aa = avec[0] * avec[0] + avec[1] * avec[1] + avec[2] * avec[2];
      double bb = bvec[0] * bvec[0] + bvec[1] * bvec[1] + bvec[2] * bvec[2];
      double ab = avec[0] * bvec[0] + avec[1] * bvec[1] + avec[2] * bvec[2];
      double theta = acos(ab / sqrt(aa * bb)), Ea = 0.0, Fa = 0.0;
      if (s.kind == ANGLE_HARMONIC) {
        double delta = theta - s.x0;
        Ea = s.half_k * delta * delta;
        Fa = s.minus_k * delta;
      }
      double factor = Fa / sqrt(aa * bb - ab * ab);
      for (int x = 0; x < 3; x++) {
        Fi[x] = ((ab / aa) * avec[x] - bvec[x]) * factor;
        Fk[x] = ((ab / bb) * bvec[x] - avec[x]) * factor;
      }
      acc[0] = Ea;
      acc[1] = (Fi[0] * avec[0] + Fk[0] * bvec[0]) + (Fi[1] * avec[1] + Fk[1] * bvec[1]) +
               (Fi[2] * avec[2] + Fk[2] * bvec[2]);
    }
    if (kspace) {
      double Rik[3];
      for (int x = 0; x < 3; x++) Rik[x] = avec[x] - bvec[x];
      double RikSq = Rik[0] * Rik[0] + Rik[1] * Rik[1] + Rik[2] * Rik[2];
      double QiQk = pt[type[si] * ntypes + type[sk]].kCoul * charge[si] * charge[sk], EL, WL;
      ew_discount(alpha, beta, RikSq, QiQk, EL, WL);
      acc[2] = EL;
      for (int x = 0; x < 3; x++) Fik[x] = WL * Rik[x] / RikSq;
    }
    double* o = aF + 9 * (size_t)m;
    for (int x = 0; x < 3; x++) {
      o[x] = Fi[x] + Fik[x];
      o[3 + x] = Fk[x] - Fik[x];
      o[6 + x] = -(Fi[x] + Fk[x]);
    }
  }
  block_sum<3>(acc);
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; k++) part[3 * blockIdx.x + k] = acc[k];
}

// EmDeeData.f90:554-640 (declared in the reference, not called by its force
// driver), in real units: z along r_kj, x the part of r_ij normal to it, y =
// z x x, phi = atan2(r_lk . y, r_lk . x); Fd = -dE/dphi.  dF per dihedral:
// force on i, j, k, l, with j's force -(Fi + Fk + Fl) so that the four sum to
// zero (the reference's line :604 adds +(Fi + Fk + Fl)).  Out per block:
// {Edihedral, Wdihedral}.
__device__ __forceinline__ void cross3(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

__global__ void __launch_bounds__(BLOCK) k_dihedrals(double* __restrict__ dF, double* __restrict__ part,
                                                     const DevDihedral* __restrict__ dih, const double* __restrict__ R,
                                                     const int* __restrict__ iperm, double L, int nd, int Npad) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[2] = { 0.0, 0.0 };
  if (m < nd) {
    DevDihedral s = dih[m];
    int si = iperm[s.i], sj = iperm[s.j], sk = iperm[s.k], sl = iperm[s.l];
    double rij[3], rkj[3], rlk[3], x[3], y[3], z[3], u[3], v[3], w[3], c[3], Fi[3], Fk[3], Fl[3];
    for (int q = 0; q < 3; q++) {
      double Rj = R[q * (size_t)Npad + sj] / L, Rk = R[q * (size_t)Npad + sk] / L;
      rij[q] = R[q * (size_t)Npad + si] / L - Rj;
      rkj[q] = Rk - Rj;
      rlk[q] = R[q * (size_t)Npad + sl] / L - Rk;
      rij[q] = L * (rij[q] - round(rij[q]));
      rkj[q] = L * (rkj[q] - round(rkj[q]));
      rlk[q] = L * (rlk[q] - round(rlk[q]));
    }
    double normRkj = sqrt(dot3(rkj, rkj));
    for (int q = 0; q < 3; q++) z[q] = rkj[q] / normRkj;
    double rz = dot3(rij, z);
    for (int q = 0; q < 3; q++) x[q] = rij[q] - rz * z[q];
    double normX = sqrt(dot3(x, x));
    for (int q = 0; q < 3; q++) x[q] = x[q] / normX;
    cross3(z, x, y);
    double a = dot3(x, rlk), b = dot3(y, rlk);
    double phi = atan2(b, a), Ed = 0.0, Fd = 0.0;
    if (s.kind == DIHEDRAL_HARMONIC) {
      const double twoPi = 6.28318530717958648;
      double delta = phi - s.x0;
      delta = delta - twoPi * round(delta / twoPi);
      Ed = s.half_k * delta * delta;
      Fd = s.minus_k * delta;
    }
    Fd = Fd / (a * a + b * b);
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
    double* o = dF + 12 * (size_t)m;
    for (int q = 0; q < 3; q++) {
      o[q] = Fi[q];
      o[3 + q] = -(Fi[q] + Fk[q] + Fl[q]);
      o[6 + q] = Fk[q];
      o[9 + q] = Fl[q];
    }
    acc[0] = Ed;
    acc[1] = (Fi[0] * rij[0] + Fk[0] * rkj[0] + Fl[0] * (rlk[0] + rkj[0])) +
             (Fi[1] * rij[1] + Fk[1] * rkj[1] + Fl[1] * (rlk[1] + rkj[1])) +
             (Fi[2] * rij[2] + Fk[2] * rkj[2] + Fl[2] * (rlk[2] + rkj[2]));
  }
  block_sum<2>(acc);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = acc[0];
    part[2 * blockIdx.x + 1] = acc[1];
  }
}

// F(slot) += contributions of the structures of atom perm[slot], in the
// fixed order of the incidence list.
__global__ void k_structs_gather(double* __restrict__ F, const double* __restrict__ bF, const double* __restrict__ aF,
                                 const double* __restrict__ dF, const int* __restrict__ incOff,
                                 const int* __restrict__ incItem,
                                 const int* __restrict__ perm, int N, int Npad) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  int a = perm[s];
  int e0 = incOff[a], e1 = incOff[a + 1];
  if (e0 == e1) return;
  double f[3] = { 0.0, 0.0, 0.0 };
  for (int e = e0; e < e1; e++) {
    int it = incItem[e], idx = it >> 4, role = it & 15;
    const double* src;
    double sign = 1.0;
    switch (role) {
      case ROLE_BOND_I: src = bF + 3 * (size_t)idx; break;
      case ROLE_BOND_J: src = bF + 3 * (size_t)idx; sign = -1.0; break;
      case ROLE_ANGLE_I: src = aF + 9 * (size_t)idx; break;
      case ROLE_ANGLE_K: src = aF + 9 * (size_t)idx + 3; break;
      case ROLE_ANGLE_J: src = aF + 9 * (size_t)idx + 6; break;
      default: src = dF + 12 * (size_t)idx + 3 * (role - ROLE_DIH_I); break;
    }
    for (int x = 0; x < 3; x++) f[x] += sign * src[x];
  }
  for (int x = 0; x < 3; x++) F[x * (size_t)Npad + s] += f[x];
}

void dev_structs_setup(Device* d, const std::vector<DevStruct>& bonds, const std::vector<DevStruct>& angles,
                       const std::vector<DevDihedral>& dihedrals) {
  if (!d->structs) d->structs = new StructState();
  StructState* st = d->structs;
  int N = d->N;
  st->nb = (int)bonds.size();
  st->na = (int)angles.size();
  st->nd = (int)dihedrals.size();
  for (const DevStruct& b : bonds)
    if (b.i < 0 || b.i >= N || b.j < 0 || b.j >= N) fatal("bonded set-up", "atom index out of range");
  for (const DevStruct& a : angles)
    if (a.i < 0 || a.i >= N || a.j < 0 || a.j >= N || a.k < 0 || a.k >= N) fatal("bonded set-up", "atom index out of range");
  for (const DevDihedral& h : dihedrals)
    if (h.i < 0 || h.i >= N || h.j < 0 || h.j >= N || h.k < 0 || h.k >= N || h.l < 0 || h.l >= N)
      fatal("bonded set-up", "atom index out of range");
  if (st->nb >= (1 << 26) || st->na >= (1 << 26) || st->nd >= (1 << 26))
    fatal("bonded set-up", "too many bonds, angles or dihedrals");
  // incidences in structure order: bonds, then angles, then dihedrals
  std::vector<std::vector<int>> inc(N);
  for (int b = 0; b < st->nb; b++) {
    inc[bonds[b].i].push_back((b << 4) | ROLE_BOND_I);
    inc[bonds[b].j].push_back((b << 4) | ROLE_BOND_J);
  }
  for (int m = 0; m < st->na; m++) {
    inc[angles[m].i].push_back((m << 4) | ROLE_ANGLE_I);
    inc[angles[m].k].push_back((m << 4) | ROLE_ANGLE_K);
    inc[angles[m].j].push_back((m << 4) | ROLE_ANGLE_J);
  }
  for (int m = 0; m < st->nd; m++) {
    inc[dihedrals[m].i].push_back((m << 4) | ROLE_DIH_I);
    inc[dihedrals[m].j].push_back((m << 4) | ROLE_DIH_J);
    inc[dihedrals[m].k].push_back((m << 4) | ROLE_DIH_K);
    inc[dihedrals[m].l].push_back((m << 4) | ROLE_DIH_L);
  }
  std::vector<int> off(N + 1, 0), items;
  for (int a = 0; a < N; a++) { items.insert(items.end(), inc[a].begin(), inc[a].end()); off[a + 1] = (int)items.size(); }
  auto put = [&](auto& buf, const auto& host) {
    buf.alloc(host.size());
    if (!host.empty()) HIP_CHECK(hipMemcpy(buf.p, host.data(), host.size() * sizeof(host[0]), hipMemcpyHostToDevice));
  };
  put(st->bonds, bonds);
  put(st->angles, angles);
  put(st->dihedrals, dihedrals);
  put(st->incOff, off);
  put(st->incItem, items);
  st->bF.alloc(3 * (size_t)st->nb);
  st->aF.alloc(9 * (size_t)st->na);
  st->dF.alloc(12 * (size_t)st->nd);
  st->part.alloc(3 * (size_t)std::max(std::max(div_up(st->nb, BLOCK), div_up(st->na, BLOCK)), div_up(st->nd, BLOCK)));
  st->result.alloc(8);
}

void dev_structs_compute(Device* d, int layer, double L, int ntypes, bool bonded, bool kspace, double alpha,
                         double beta, double* out) {
  for (int k = 0; k < 7; k++) out[k] = 0.0;
  StructState* st = d->structs;
  if (!st || (st->nb == 0 && st->na == 0 && st->nd == 0) || d->N == 0) return;
  int Np = d->Npad;
  double* F = d->F.p + 3 * (size_t)Np * layer;
  const DevPair* pt = d->pairs.p + (size_t)layer * ntypes * ntypes;
  int gb = div_up(st->nb, BLOCK), ga = div_up(st->na, BLOCK), gd = div_up(st->nd, BLOCK);
  if (st->nb > 0) {
    hipLaunchKernelGGL(k_bonds, dim3(gb), dim3(BLOCK), 0, d->stream, st->bF.p, st->part.p, st->bonds.p, d->R.p,
                       d->iperm.p, d->type.p, d->charge.p, pt, ntypes, L, (int)bonded, (int)kspace, alpha, beta,
                       st->nb, Np);
    launch_check("bonds");
    hipLaunchKernelGGL(k_reduce_partials<3>, dim3(1), dim3(BLOCK), 0, d->stream, st->result.p, st->part.p, gb);
  }
  if (st->na > 0) {
    // the angle kernel reuses the partials buffer after the bond reduction (same stream)
    hipLaunchKernelGGL(k_angles, dim3(ga), dim3(BLOCK), 0, d->stream, st->aF.p, st->part.p, st->angles.p, d->R.p,
                       d->iperm.p, d->type.p, d->charge.p, pt, ntypes, L, (int)bonded, (int)kspace, alpha, beta,
                       st->na, Np);
    launch_check("angles");
    hipLaunchKernelGGL(k_reduce_partials<3>, dim3(1), dim3(BLOCK), 0, d->stream, st->result.p + 3, st->part.p, ga);
  }
  // dihedrals are bonded terms only (no k-space discount of their 1-4 pairs)
  if (st->nd > 0 && bonded) {
    hipLaunchKernelGGL(k_dihedrals, dim3(gd), dim3(BLOCK), 0, d->stream, st->dF.p, st->part.p, st->dihedrals.p, d->R.p,
                       d->iperm.p, L, st->nd, Np);
    launch_check("dihedrals");
    hipLaunchKernelGGL(k_reduce_partials<2>, dim3(1), dim3(BLOCK), 0, d->stream, st->result.p + 6, st->part.p, gd);
  } else if (st->nd > 0) {
    HIP_CHECK(hipMemsetAsync(st->dF.p, 0, sizeof(double) * 12 * (size_t)st->nd, d->stream));
  }
  hipLaunchKernelGGL(k_structs_gather, dim3(div_up(d->N, BLOCK)), dim3(BLOCK), 0, d->stream, F, st->bF.p, st->aF.p,
                     st->dF.p, st->incOff.p, st->incItem.p, d->perm.p, d->N, Np);
  launch_check("bonded gather");
  HIP_CHECK(hipMemcpyAsync(d->hresult, st->result.p, 8 * sizeof(double), hipMemcpyDeviceToHost, d->stream));
  HIP_CHECK(hipStreamSynchronize(d->stream));
  double* h = d->hresult;
  if (st->nb == 0) h[0] = h[1] = h[2] = 0.0;
  if (st->na == 0) h[3] = h[4] = h[5] = 0.0;
  if (st->nd == 0 || !bonded) h[6] = h[7] = 0.0;
  out[0] = h[0];
  out[1] = h[1];
  out[2] = h[3];
  out[3] = h[4];
  out[4] = h[2] + h[5];
  out[5] = h[6];
  out[6] = h[7];
}

}  // namespace emdee
