// hmc.hip -- batched hybrid Monte Carlo of many independent Lennard-Jones
// replicas in one persistent launch.
//
// There is no counterpart in the reference: a reference user runs replicas
// one after another, each a tEmDee driven by EmDee_upload("momenta"),
// EmDee_boost / EmDee_displace / EmDee_boost (velocity Verlet, mass 1,
// pair_lj_cut) and a Metropolis test on the total energy.  Here one
// workgroup owns one replica for a whole HMC cycle (momentum refresh, an
// nsteps-long trajectory, accept/reject, restore on rejection), and one
// kernel launch runs several cycles for every replica; the launch is
// captured once in a HIP graph and replayed.
//
// Per replica (N <= 8192 atoms) the workgroup keeps, in LDS, the 32-bit
// fixed-point coordinates of the current step, so all pair gathers of the
// mixed form are LDS reads.  Pairs come from a Verlet list per replica in
// global memory (16-bit indices, column-major; kept between launches with
// its build coordinates): built from a cell list
// (counting sort into cells of side >= (rc + skin)/DIV, stencil of
// (2 DIV + 1)^3 cells) and rebuilt before a force evaluation whenever some
// atom has moved more than skin/2 since the last build -- the rule of
// neighbor_lists.f90:41-59 in its simplest sufficient form -- which also
// covers the coordinates restored after a rejected move.  A replica whose
// list would overflow its capacity (denser than LJ close packing) takes every
// atom as a candidate instead.  The pair set is the reference's: r^2 < rc^2 on the FP64
// scaled coordinates (neighbor_lists.f90:244), with FP32 candidates in the
// band around the cutoff re-tested in FP64.  Pair terms are FP32 (mixed form)
// or FP64 (plain form), accumulated per atom in FP64.
//
// Random numbers come from a counter-based generator (splitmix64 finaliser,
// Steele, Lea & Flood 2014) keyed by (seed, replica, cycle, atom, component),
// so the momenta and uniforms do not depend on the launch shape; Gaussian
// momenta use the Box-Muller transform.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <numeric>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/emdee_gpu.h"
#include "device_state.h"

namespace emdee {

namespace {
constexpr int HMC_NMAX = 8192;   // atoms per replica
constexpr int HMC_CMAX = 6144;   // cells per replica
constexpr int HMC_HIST = 1024;   // cycles of potential-energy history
const double TWO_PI_D = 6.283185307179586;

struct ReplicaDesc {
  int off, N, M, Ml;       // atom offset and count; cells per edge for the force walk and for the list build
  double L, kT;
  float K, lo, hi, fpad;   // fixed-point cutoff band (mixed form)
  double Rc2s, unit;       // FP64 scaled cutoff; L / 2^32
  double Rl2s, trig2s;     // list radius (rc + skin)^2 and (skin/2)^2, scaled units^2
};
constexpr double HMC_SKIN = 0.3;   // list skin (LJ units; 0.2 and 0.45 measured slower)

__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
// uniform in (0, 1) for (seed, stream, counter)
__host__ __device__ inline double hmc_uniform(uint64_t seed, uint64_t stream, uint64_t counter) {
  uint64_t x = mix64(seed ^ mix64(stream * 0x9E3779B97F4A7C15ULL + counter));
  return ((double)(x >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
__host__ __device__ inline uint64_t hmc_stream(int replica, int cycle) {
  return ((uint64_t)(uint32_t)replica << 32) | (uint32_t)cycle;
}
constexpr uint64_t HMC_ACCEPT_COUNTER = 1ULL << 40;

__device__ __forceinline__ double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

template <int BS>
__device__ double block_total(double v, double* red) {
  v = wave_sum(v);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int k = 0; k < BS / 64; k++) s += red[k];
  return s;
}
}  // namespace

// One HMC workgroup per replica.  mode 0: forces and potential energy of the
// current coordinates only; mode 1: ncycles full cycles.
template <int BS, int DIV, bool MIXED>
__global__ void __launch_bounds__(BS)
k_hmc(const ReplicaDesc* __restrict__ desc, const int* __restrict__ order, double4* __restrict__ R,
      double4* __restrict__ P, double4* __restrict__ F, double4* __restrict__ Rsave, double4* __restrict__ Fsave,
      double* __restrict__ U, int* __restrict__ cycle, long long* __restrict__ accepted,
      double* __restrict__ lastDH, double rc, double dt, int nsteps, uint64_t seed, int ncycles, int mode,
      unsigned short* __restrict__ nlist, unsigned short* __restrict__ nnb, uint4* __restrict__ qbuild, int cap,
      int* __restrict__ lstate, double* __restrict__ Uhist, int histLen) {
  __shared__ unsigned int qx[HMC_NMAX], qy[HMC_NMAX], qz[HMC_NMAX];
  __shared__ unsigned short clist[HMC_NMAX];
  __shared__ int cstart[HMC_CMAX + 1];
  __shared__ unsigned int ccount2[HMC_CMAX / 2];   // 16-bit cell counts, two per word
  __shared__ double red[BS / 64];
  __shared__ int scanTmp[BS];
  __shared__ int cycleNow;
  __shared__ double Unow;
  __shared__ int listState;   // 0: no valid list, 1: valid list, 2: overflow (cell walk)

  const int r = order[blockIdx.x];
  const ReplicaDesc d = desc[r];
  const int N = d.N;
  double4* Rr = R + d.off;
  double4* Pr = P + d.off;
  double4* Fr = F + d.off;
  const double L = d.L, hdt = 0.5 * dt;
  const float c0 = 4.0f, c1 = 24.0f, c2 = 1.0f;   // pair_lj_cut with epsilon = sigma = 1

  // cell list of the current coordinates (fixed point in LDS)
  auto count_of = [&](int c) { return (int)((ccount2[c >> 1] >> ((c & 1) * 16)) & 0xffffu); };
  auto bump = [&](int c) {   // returns the previous count of cell c
    unsigned int old = atomicAdd(&ccount2[c >> 1], 1u << ((c & 1) * 16));
    return (int)((old >> ((c & 1) * 16)) & 0xffffu);
  };
  // fixed-point coordinates of the current R into LDS; returns the largest
  // squared displacement since the last list build (scaled units^2)
  auto refresh = [&]() -> double {
    double m = 0.0;
    for (int i = threadIdx.x; i < N; i += BS) {
      double4 ri = Rr[i];
      double s[3] = { ri.x / L, ri.y / L, ri.z / L };
      unsigned int u[3];
      for (int x = 0; x < 3; x++) {
        double f = s[x] - floor(s[x]);
        u[x] = (unsigned int)(unsigned long long)llrint(f * 4294967296.0);
      }
      qx[i] = u[0]; qy[i] = u[1]; qz[i] = u[2];
      uint4 b = qbuild[d.off + i];
      double ex = (double)(int)(u[0] - b.x), ey = (double)(int)(u[1] - b.y), ez = (double)(int)(u[2] - b.z);
      m = fmax(m, (ex * ex + ey * ey + ez * ez) * (1.0 / 18446744073709551616.0));
    }
    m = wave_max(m);
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[w] = m;
    __syncthreads();
    double t = 0.0;
    for (int k = 0; k < BS / 64; k++) t = fmax(t, red[k]);
    return t;
  };
  // cell list of the coordinates in LDS, cells of side L / Mc
  auto build_cells = [&](int Mc) {
    const int ncc = Mc * Mc * Mc;
    for (int c = threadIdx.x; c < (ncc + 1) / 2; c += BS) ccount2[c] = 0;
    __syncthreads();
    auto cell_of = [&](int i) {
      return (int)(((unsigned long long)qx[i] * Mc) >> 32) +
             Mc * ((int)(((unsigned long long)qy[i] * Mc) >> 32) + Mc * (int)(((unsigned long long)qz[i] * Mc) >> 32));
    };
    for (int i = threadIdx.x; i < N; i += BS) bump(cell_of(i));
    __syncthreads();
    // exclusive scan of the counts: per-thread chunk sums, block scan, chunks
    int chunk = (ncc + BS - 1) / BS, c0i = threadIdx.x * chunk, c1i = min(ncc, c0i + chunk);
    int local = 0;
    for (int c = c0i; c < c1i; c++) local += count_of(c);
    scanTmp[threadIdx.x] = local;
    __syncthreads();
    for (int off = 1; off < BS; off <<= 1) {
      int v = threadIdx.x >= off ? scanTmp[threadIdx.x - off] : 0;
      __syncthreads();
      scanTmp[threadIdx.x] += v;
      __syncthreads();
    }
    int run = scanTmp[threadIdx.x] - local;
    for (int c = c0i; c < c1i; c++) { cstart[c] = run; run += count_of(c); }
    if (threadIdx.x == 0) cstart[ncc] = N;
    __syncthreads();
    for (int c = threadIdx.x; c < (ncc + 1) / 2; c += BS) ccount2[c] = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < N; i += BS) {
      int c = cell_of(i);
      clist[cstart[c] + bump(c)] = (unsigned short)i;
    }
    __syncthreads();
  };
  // visits, for atom i, every atom j of the cells within Mc-cell reach that
  // can lie within sqrt(R2s) (scaled units) of it, culled row by row; the
  // atom itself is among them.  Mc < 2 DIV + 1: one cell holds every atom.
  auto walk = [&](int i, int Mc, double R2s, auto&& visit) {
    const int w = (Mc >= 2 * DIV + 1) ? DIV : 0;
    // cutoff^2 in cell units, with margins far above the FP32 rounding of the
    // culling (which can then only keep extra cells)
    const float rcell2 = (float)(R2s * (double)Mc * (double)Mc * (1.0 + 1e-5) + 1e-5);
    unsigned int ui = qx[i], vi = qy[i], wi = qz[i];
    unsigned long long px = (unsigned long long)ui * Mc, py = (unsigned long long)vi * Mc,
                       pz = (unsigned long long)wi * Mc;
    int cx = (int)(px >> 32), cy = (int)(py >> 32), cz = (int)(pz >> 32);
    // position of atom i inside its cell, in cell units
    const float ox = (float)(unsigned int)px * 2.3283064e-10f, oy = (float)(unsigned int)py * 2.3283064e-10f,
                oz = (float)(unsigned int)pz * 2.3283064e-10f;
    // cells are numbered x fastest, so the cells x-w..x+w of one (y, z) row
    // are one contiguous slot range (two where the row wraps, visited in
    // the order of increasing dx); w = 0 takes every atom at once
    for (int dz = -w; dz <= w; dz++) {
      int z = cz + dz; z = z < 0 ? z + Mc : (z >= Mc ? z - Mc : z);
      float gz = fmaxf(dz > 0 ? dz - oz : (dz < 0 ? oz - dz - 1 : 0.0f), 0.0f);
      for (int dy = -w; dy <= w; dy++) {
        int y = cy + dy; y = y < 0 ? y + Mc : (y >= Mc ? y - Mc : y);
        int base = Mc * (y + Mc * z), xa = cx - w, xb = cx + w;
        if (w > 0) {   // rows and cells beyond the cutoff sphere hold no partner
          float gy = fmaxf(dy > 0 ? dy - oy : (dy < 0 ? oy - dy - 1 : 0.0f), 0.0f);
          float rem = rcell2 - gy * gy - gz * gz;
          if (rem <= 0.0f) continue;
          float ex = sqrtf(rem) + 1e-4f;
          xa = cx + max((int)floorf(ox - ex), -w);
          xb = cx + min((int)floorf(ox + ex), w);
        }
        int r0a, r0b, r1a = 0, r1b = -1;   // cell ranges [r0a, r0b] then [r1a, r1b]
        if (xa < 0) { r0a = xa + Mc; r0b = Mc - 1; r1a = 0; r1b = xb; }
        else if (xb >= Mc) { r0a = xa; r0b = Mc - 1; r1a = 0; r1b = xb - Mc; }
        else { r0a = xa; r0b = xb; }
        for (int part = 0; part < 2; part++) {
          int ca = part == 0 ? r0a : r1a, cb = part == 0 ? r0b : r1b;
          if (cb < ca) continue;
          int t0 = (w == 0) ? 0 : cstart[base + ca], t1 = (w == 0) ? N : cstart[base + cb + 1];
          for (int t = t0; t < t1; t++) visit((w == 0) ? t : (int)clist[t]);
        }
      }
    }
  };
  // Verlet list at radius sqrt(d.Rl2s) from the coordinates in LDS (every
  // atom within it, in walk order; the FP32 test keeps a margin, so the list
  // may hold a few more); the build coordinates are kept for the trigger
  auto build_list = [&]() {
    build_cells(d.Ml);
    const float rl2 = (float)(d.Rl2s * 18446744073709551616.0 * (1.0 + 1e-4));
    int over = 0;
    for (int i = threadIdx.x; i < N; i += BS) {
      const unsigned int ui = qx[i], vi = qy[i], wi = qz[i];
      int n = 0;
      walk(i, d.Ml, d.Rl2s, [&](int j) {
        float ddx = (float)(int)(ui - qx[j]), ddy = (float)(int)(vi - qy[j]), ddz = (float)(int)(wi - qz[j]);
        float r2 = ddx * ddx + ddy * ddy + ddz * ddz;
        if (r2 < rl2 && j != i) {
          if (n < cap) nlist[(size_t)d.off * cap + (size_t)n * N + i] = (unsigned short)j;
          n++;
        }
      });
      nnb[d.off + i] = (unsigned short)min(n, cap);
      qbuild[d.off + i] = make_uint4(ui, vi, wi, 0u);
      over |= n > cap;
    }
    // any overflow: this replica walks the cells from now on
    over = __syncthreads_or(over);
    if (threadIdx.x == 0) listState = over ? 2 : 1;
    __syncthreads();
  };

  // forces F = -dU/dR and, with E_ true, the potential energy of the replica
  // (only the last step of a trajectory needs it)
  auto forces = [&](auto E_) -> double {
    constexpr bool WANT_E = decltype(E_)::value;
    // the list: rebuilt when it is missing or an atom moved more than skin/2
    const double moved = refresh();
    if (listState != 2 && (listState == 0 || moved > d.trig2s)) build_list();
    const bool useList = listState == 1;
    double e = 0.0;
    for (int i = threadIdx.x; i < N; i += BS) {
      const unsigned int ui = qx[i], vi = qy[i], wi = qz[i];
      double4 ri = Rr[i];
      double fx = 0, fy = 0, fz = 0, ei = 0;
      auto pair = [&](int j) {
        if (MIXED) {
          // the atom itself (zero separation) is masked, not branched over
          float ddx = (float)(int)(ui - qx[j]), ddy = (float)(int)(vi - qy[j]), ddz = (float)(int)(wi - qz[j]);
          float r2 = ddx * ddx + ddy * ddy + ddz * ddz;
          bool in = r2 < d.lo && j != i;
          if (r2 >= d.lo && r2 < d.hi) {
            double4 rj = Rr[j];
            double ex = ri.x / L - rj.x / L, ey = ri.y / L - rj.y / L, ez = ri.z / L - rj.z / L;
            ex = ex - rint(ex); ey = ey - rint(ey); ez = ez - rint(ez);
            double q2;
            {
#pragma clang fp contract(off)
              q2 = ex * ex + ey * ey + ez * ez;
            }
            in = q2 < d.Rc2s;
          }
          if (in) {
            float invR2 = d.K * __builtin_amdgcn_rcpf(r2);
            float sr2 = c2 * invR2, sr6 = sr2 * sr2 * sr2;
            float W = c1 * sr6 * (sr6 + sr6 - 1.0f);
            float g = W * invR2;
            fx += (double)(g * ddx);
            fy += (double)(g * ddy);
            fz += (double)(g * ddz);
            if (WANT_E) ei += (double)(c0 * sr6 * (sr6 - 1.0f));
          }
        } else {
          if (j == i) return;
          double4 rj = Rr[j];
          double ex = ri.x / L - rj.x / L, ey = ri.y / L - rj.y / L, ez = ri.z / L - rj.z / L;
          ex = ex - rint(ex); ey = ey - rint(ey); ez = ez - rint(ez);
          double q2;
          {
#pragma clang fp contract(off)
            q2 = ex * ex + ey * ey + ez * ez;
          }
          if (q2 < d.Rc2s) {
            double invR2 = 1.0 / (L * L * q2);
            double sr2 = invR2, sr6 = sr2 * sr2 * sr2, sr12 = sr6 * sr6;
            double W = 24.0 * (sr12 + sr12 - sr6);
            double g = W * invR2;
            fx += g * ex;
            fy += g * ey;
            fz += g * ez;
            if (WANT_E) ei += 4.0 * (sr12 - sr6);
          }
        }
      };
      if (useList) {
        // eight indices loaded before their pairs are evaluated, so that
        // the loads overlap; past the list end the atom itself stands in,
        // which the pair body skips (FP64) or masks (mixed), so the eight
        // are evaluated without a branch
        const unsigned short* __restrict__ li = nlist + (size_t)d.off * cap + i;
        const int n = nnb[d.off + i];
        for (int k = 0; k < n; k += 8) {
          int jb[8];
#pragma unroll
          for (int u = 0; u < 8; u++) jb[u] = k + u < n ? (int)li[(size_t)(k + u) * N] : i;
#pragma unroll
          for (int u = 0; u < 8; u++) pair(jb[u]);
        }
      } else {
        // a list that would overflow its capacity (more neighbours within
        // rc + skin than LJ atoms at close packing hold): every atom
        for (int j = 0; j < N; j++) pair(j);
      }
      double sc = MIXED ? d.unit : L;
      Fr[i] = make_double4(sc * fx, sc * fy, sc * fz, 0.0);
      e += ei;
    }
    return WANT_E ? 0.5 * block_total<BS>(e, red) : 0.0;
  };

  // the list and its build coordinates persist between launches (written
  // together), so a list still valid for the current coordinates is kept
  if (threadIdx.x == 0) listState = lstate[r];
  __syncthreads();
  if (mode == 0) {
    double u = forces(std::true_type());
    if (threadIdx.x == 0) { U[r] = u; lstate[r] = listState; }
    return;
  }
  const double sq = sqrt(d.kT);
  // per-replica scalars live in LDS for the whole launch
  if (threadIdx.x == 0) { cycleNow = cycle[r]; Unow = U[r]; }
  __syncthreads();
  for (int cyc = 0; cyc < ncycles; cyc++) {
    const int cnum = cycleNow;
    const uint64_t stream = hmc_stream(r, cnum);
    // momentum refresh (mass 1) and saved state
    double k = 0.0;
    for (int i = threadIdx.x; i < N; i += BS) {
      double n[3];
      for (int x = 0; x < 3; x++) {
        double u1 = hmc_uniform(seed, stream, 6ULL * i + 2 * x), u2 = hmc_uniform(seed, stream, 6ULL * i + 2 * x + 1);
        n[x] = sq * (sqrt(-2.0 * log(u1)) * cos(TWO_PI_D * u2));
      }
      Pr[i] = make_double4(n[0], n[1], n[2], 0.0);
      Rsave[d.off + i] = Rr[i];
      Fsave[d.off + i] = Fr[i];
      k += n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    }
    const double U0 = Unow, H0 = U0 + 0.5 * block_total<BS>(k, red);
    double U1 = U0;
    for (int s = 0; s < nsteps; s++) {
      for (int i = threadIdx.x; i < N; i += BS) {
        double4 p = Pr[i], f = Fr[i], x = Rr[i];
        p.x = p.x + hdt * f.x; p.y = p.y + hdt * f.y; p.z = p.z + hdt * f.z;
        x.x = x.x + dt * p.x; x.y = x.y + dt * p.y; x.z = x.z + dt * p.z;
        Pr[i] = p;
        Rr[i] = x;
      }
      __syncthreads();
      if (s == nsteps - 1) U1 = forces(std::true_type());
      else forces(std::false_type());
      for (int i = threadIdx.x; i < N; i += BS) {
        double4 p = Pr[i], f = Fr[i];
        p.x = p.x + hdt * f.x; p.y = p.y + hdt * f.y; p.z = p.z + hdt * f.z;
        Pr[i] = p;
      }
      __syncthreads();
    }
    k = 0.0;
    for (int i = threadIdx.x; i < N; i += BS) {
      double4 p = Pr[i];
      k += p.x * p.x + p.y * p.y + p.z * p.z;
    }
    const double H1 = U1 + 0.5 * block_total<BS>(k, red);
    const double dH = H1 - H0;
    const bool accept = dH <= 0.0 || hmc_uniform(seed, stream, HMC_ACCEPT_COUNTER) < exp(-dH / d.kT);
    if (!accept)
      for (int i = threadIdx.x; i < N; i += BS) { Rr[i] = Rsave[d.off + i]; Fr[i] = Fsave[d.off + i]; }
    __syncthreads();
    if (threadIdx.x == 0) {
      Unow = accept ? U1 : U0;
      cycleNow = cnum + 1;
      U[r] = Unow;
      accepted[r] += accept ? 1 : 0;
      lastDH[r] = dH;
      cycle[r] = cycleNow;
      Uhist[(size_t)(cnum % histLen) * gridDim.x + r] = Unow;   // the grid holds one workgroup per replica
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) lstate[r] = listState;
}

}  // namespace emdee

// ------------------------------------------------------------ host side

using namespace emdee;

struct EmDeeHMC {
  int nrep = 0, total = 0, nsteps = 0, cyclesPerLaunch = 1, precision = 1;
  double rc = 2.5, dt = 0.005;
  uint64_t seed = 0;
  hipStream_t stream = nullptr;
  std::vector<ReplicaDesc> hdesc;
  std::vector<int> offsets, hN;
  DBuf<ReplicaDesc> desc;
  DBuf<int> order, cycle;
  DBuf<double4> R, P, F, Rsave, Fsave;
  DBuf<double> U, lastDH;
  DBuf<long long> accepted;
  int cap = 0;                   // Verlet-list entries per atom
  DBuf<unsigned short> nlist, nnb;
  DBuf<uint4> qbuild;
  DBuf<int> lstate;              // per replica: 0 no list, 1 valid list, 2 overflow
  DBuf<double> Uhist;            // potential energy after each of the last HMC_HIST cycles [cycle % HMC_HIST][replica]
  int variant = -1;            // index into the variant table
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double tuneMs[6] = { 0, 0, 0, 0, 0, 0 };
};

namespace {
using HmcKernel = void (*)(const ReplicaDesc*, const int*, double4*, double4*, double4*, double4*, double4*, double*,
                           int*, long long*, double*, double, double, int, uint64_t, int, int, unsigned short*,
                           unsigned short*, uint4*, int, int*, double*, int);
struct Variant { int bs, div; HmcKernel mixed, plain; };
const Variant VARIANTS[] = {
  { 256, 1, k_hmc<256, 1, true>, k_hmc<256, 1, false> },
  { 512, 1, k_hmc<512, 1, true>, k_hmc<512, 1, false> },
  { 1024, 1, k_hmc<1024, 1, true>, k_hmc<1024, 1, false> },
  { 256, 2, k_hmc<256, 2, true>, k_hmc<256, 2, false> },
  { 512, 2, k_hmc<512, 2, true>, k_hmc<512, 2, false> },
  { 1024, 2, k_hmc<1024, 2, true>, k_hmc<1024, 2, false> },
};
constexpr int NVARIANTS = 6;

bool variant_fits(const EmDeeHMC* h, int v) {
  for (const ReplicaDesc& d : h->hdesc) {
    int M = std::max(1, (int)std::floor(VARIANTS[v].div * d.L / h->rc));
    if ((long)M * M * M > HMC_CMAX) return false;
  }
  return true;
}

void set_cells(EmDeeHMC* h, int div) {
  for (ReplicaDesc& d : h->hdesc) {
    int M = std::max(1, (int)std::floor(div * d.L / h->rc));
    d.M = M >= 2 * div + 1 ? M : 1;
    int Ml = std::max(1, (int)std::floor(div * d.L / (h->rc + HMC_SKIN)));
    d.Ml = Ml >= 2 * div + 1 ? Ml : 1;
  }
  HIP_CHECK(hipMemcpy(h->desc.p, h->hdesc.data(), sizeof(ReplicaDesc) * h->nrep, hipMemcpyHostToDevice));
}

void launch(EmDeeHMC* h, int v, int ncycles, int mode) {
  const Variant& var = VARIANTS[v];
  HmcKernel k = h->precision ? var.mixed : var.plain;
  hipLaunchKernelGGL(k, dim3(h->nrep), dim3(var.bs), 0, h->stream, h->desc.p, h->order.p, h->R.p, h->P.p, h->F.p,
                     h->Rsave.p, h->Fsave.p, h->U.p, h->cycle.p, h->accepted.p, h->lastDH.p, h->rc, h->dt, h->nsteps,
                     h->seed, ncycles, mode, h->nlist.p, h->nnb.p, h->qbuild.p, h->cap, h->lstate.p, h->Uhist.p,
                     HMC_HIST);
  launch_check("hmc");
}
}  // namespace

extern "C" {

EmDeeHMC* EmDeeHMC_create(int nrep, const int* natoms, const double* box, const double* kT, const double* R,
                          double rc, double dt, int nsteps, unsigned long long seed, int precision) {
  const char* task = "batched HMC set-up";
  if (nrep < 1) fatal(task, "no replicas");
  EmDeeHMC* h = new EmDeeHMC();
  h->nrep = nrep; h->rc = rc; h->dt = dt; h->nsteps = nsteps; h->seed = seed; h->precision = precision ? 1 : 0;
  HIP_CHECK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  h->offsets.resize(nrep);
  h->hN.assign(natoms, natoms + nrep);
  h->hdesc.resize(nrep);
  int off = 0;
  for (int r = 0; r < nrep; r++) {
    if (natoms[r] < 2 || natoms[r] > HMC_NMAX) fatal(task, "replica sizes must be between 2 and 8192 atoms");
    if (box[r] < 2.0 * rc) fatal(task, "replica box shorter than twice the cutoff");
    ReplicaDesc& d = h->hdesc[r];
    d.off = off; d.N = natoms[r]; d.M = 1; d.L = box[r]; d.kT = kT[r];
    double two32 = 4294967296.0, rc2s = (rc * rc) / (box[r] * box[r]), rc2u = rc2s * two32 * two32;
    d.K = (float)(two32 * two32 / (box[r] * box[r]));
    // FP32 rounding plus the quantisation of the fixed-point coordinates
    const double q = 1.0 / two32, band = 2e-6 + 4.0 * q / std::sqrt(rc2s) + 3.0 * q * q / rc2s;
    d.lo = (float)(rc2u * (1.0 - band));
    d.hi = (float)(rc2u * (1.0 + band));
    d.Rc2s = rc2s;
    d.unit = box[r] / two32;
    const double rl = rc + HMC_SKIN;
    d.Rl2s = (rl * rl) / (box[r] * box[r]);
    d.trig2s = (0.25 * HMC_SKIN * HMC_SKIN) / (box[r] * box[r]);
    h->offsets[r] = off;
    off += natoms[r];
  }
  h->total = off;
  // list capacity: 1.6 times the mean count at the densest replica's
  // density, plus room for fluctuations (a replica that overflows walks its
  // cells instead)
  {
    double rho = 0.0, rl = rc + HMC_SKIN;
    for (int r = 0; r < nrep; r++) rho = std::max(rho, natoms[r] / (box[r] * box[r] * box[r]));
    int cap = (int)(1.6 * 4.18879020478639 * rl * rl * rl * rho) + 24;
    h->cap = std::min(((cap + 7) / 8) * 8, HMC_NMAX);
  }
  // largest replicas first: the dispatcher hands out workgroups in order
  std::vector<int> ord(nrep);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return natoms[a] > natoms[b]; });
  h->desc.alloc(nrep); h->order.alloc(nrep); h->cycle.alloc(nrep); h->U.alloc(nrep); h->lastDH.alloc(nrep);
  h->accepted.alloc(nrep);
  for (DBuf<double4>* b : { &h->R, &h->P, &h->F, &h->Rsave, &h->Fsave }) b->alloc(off);
  h->nlist.alloc((size_t)off * h->cap); h->nnb.alloc(off); h->qbuild.alloc(off); h->lstate.alloc(nrep);
  h->Uhist.alloc((size_t)HMC_HIST * nrep);
  std::vector<double4> R4(off);
  for (int i = 0; i < off; i++) R4[i] = make_double4(R[3 * i], R[3 * i + 1], R[3 * i + 2], 0.0);
  HIP_CHECK(hipMemcpy(h->R.p, R4.data(), sizeof(double4) * off, hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(h->order.p, ord.data(), sizeof(int) * nrep, hipMemcpyHostToDevice));
  std::vector<int> zeroI(nrep, 0);
  std::vector<long long> zeroL(nrep, 0);
  HIP_CHECK(hipMemcpy(h->cycle.p, zeroI.data(), sizeof(int) * nrep, hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(h->lstate.p, zeroI.data(), sizeof(int) * nrep, hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(h->accepted.p, zeroL.data(), sizeof(long long) * nrep, hipMemcpyHostToDevice));
  // default variant: the first that fits; forces of the initial coordinates
  for (int v = 0; v < NVARIANTS; v++)
    if (variant_fits(h, v)) { h->variant = v; break; }
  if (h->variant < 0) fatal(task, "replica boxes too large for the cell table");
  set_cells(h, VARIANTS[h->variant].div);
  launch(h, h->variant, 0, 0);
  HIP_CHECK(hipStreamSynchronize(h->stream));
  return h;
}

// Autotuner: times one cycle of every variant that fits (block size x cell
// division) from the same state, restores the state, and keeps the fastest.
int EmDeeHMC_tune(EmDeeHMC* h) {
  int n = h->total, nr = h->nrep;
  DBuf<double4> R0, F0;
  DBuf<double> U0;
  DBuf<int> c0;
  DBuf<long long> a0;
  R0.alloc(n); F0.alloc(n); U0.alloc(nr); c0.alloc(nr); a0.alloc(nr);
  HIP_CHECK(hipMemcpyAsync(R0.p, h->R.p, sizeof(double4) * n, hipMemcpyDeviceToDevice, h->stream));
  HIP_CHECK(hipMemcpyAsync(F0.p, h->F.p, sizeof(double4) * n, hipMemcpyDeviceToDevice, h->stream));
  HIP_CHECK(hipMemcpyAsync(U0.p, h->U.p, sizeof(double) * nr, hipMemcpyDeviceToDevice, h->stream));
  HIP_CHECK(hipMemcpyAsync(c0.p, h->cycle.p, sizeof(int) * nr, hipMemcpyDeviceToDevice, h->stream));
  HIP_CHECK(hipMemcpyAsync(a0.p, h->accepted.p, sizeof(long long) * nr, hipMemcpyDeviceToDevice, h->stream));
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  int best = -1;
  double bestMs = 1e300;
  for (int v = 0; v < NVARIANTS; v++) {
    h->tuneMs[v] = -1.0;
    if (!variant_fits(h, v)) continue;
    set_cells(h, VARIANTS[v].div);
    HIP_CHECK(hipEventRecord(e0, h->stream));
    launch(h, v, 1, 1);
    HIP_CHECK(hipEventRecord(e1, h->stream));
    HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    h->tuneMs[v] = ms;
    if (ms < bestMs) { bestMs = ms; best = v; }
    HIP_CHECK(hipMemcpyAsync(h->R.p, R0.p, sizeof(double4) * n, hipMemcpyDeviceToDevice, h->stream));
    HIP_CHECK(hipMemcpyAsync(h->F.p, F0.p, sizeof(double4) * n, hipMemcpyDeviceToDevice, h->stream));
    HIP_CHECK(hipMemcpyAsync(h->U.p, U0.p, sizeof(double) * nr, hipMemcpyDeviceToDevice, h->stream));
    HIP_CHECK(hipMemcpyAsync(h->cycle.p, c0.p, sizeof(int) * nr, hipMemcpyDeviceToDevice, h->stream));
    HIP_CHECK(hipMemcpyAsync(h->accepted.p, a0.p, sizeof(long long) * nr, hipMemcpyDeviceToDevice, h->stream));
  }
  HIP_CHECK(hipStreamSynchronize(h->stream));
  HIP_CHECK(hipEventDestroy(e0));
  HIP_CHECK(hipEventDestroy(e1));
  R0.release(); F0.release(); U0.release(); c0.release(); a0.release();
  h->variant = best;
  set_cells(h, VARIANTS[best].div);
  if (h->exec) { (void)hipGraphExecDestroy(h->exec); h->exec = nullptr; }
  if (h->graph) { (void)hipGraphDestroy(h->graph); h->graph = nullptr; }
  return best;
}

// Runs `launches` graph launches of `cycles` HMC cycles each (the graph is
// captured on the first call and whenever the number of cycles changes).
void EmDeeHMC_run(EmDeeHMC* h, int launches, int cycles) {
  if (cycles < 1 || launches < 1) return;
  if (!h->exec || cycles != h->cyclesPerLaunch) {
    if (h->exec) { (void)hipGraphExecDestroy(h->exec); h->exec = nullptr; }
    if (h->graph) { (void)hipGraphDestroy(h->graph); h->graph = nullptr; }
    h->cyclesPerLaunch = cycles;
    HIP_CHECK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    launch(h, h->variant, cycles, 1);
    HIP_CHECK(hipStreamEndCapture(h->stream, &h->graph));
    HIP_CHECK(hipGraphInstantiate(&h->exec, h->graph, nullptr, nullptr, 0));
  }
  for (int k = 0; k < launches; k++) HIP_CHECK(hipGraphLaunch(h->exec, h->stream));
}

void EmDeeHMC_synchronize(EmDeeHMC* h) { HIP_CHECK(hipStreamSynchronize(h->stream)); }

// Coordinates (3 * total atoms, replica order), potential energies, accepted
// moves and the last energy change of every replica; any pointer may be null.
void EmDeeHMC_get(EmDeeHMC* h, double* R, double* U, long long* accepted, double* lastDH) {
  HIP_CHECK(hipStreamSynchronize(h->stream));
  if (R) {
    std::vector<double4> R4(h->total);
    HIP_CHECK(hipMemcpy(R4.data(), h->R.p, sizeof(double4) * h->total, hipMemcpyDeviceToHost));
    for (int i = 0; i < h->total; i++) { R[3 * i] = R4[i].x; R[3 * i + 1] = R4[i].y; R[3 * i + 2] = R4[i].z; }
  }
  if (U) HIP_CHECK(hipMemcpy(U, h->U.p, sizeof(double) * h->nrep, hipMemcpyDeviceToHost));
  if (accepted) HIP_CHECK(hipMemcpy(accepted, h->accepted.p, sizeof(long long) * h->nrep, hipMemcpyDeviceToHost));
  if (lastDH) HIP_CHECK(hipMemcpy(lastDH, h->lastDH.p, sizeof(double) * h->nrep, hipMemcpyDeviceToHost));
}

// Potential energies after each of the last `cycles` cycles (at most 1024
// and at most the cycles run), oldest first: U[c * nrep + replica].  Returns
// the number of cycles copied.  Every replica runs the same cycles, so the
// first replica's cycle counter numbers them.
int EmDeeHMC_history(EmDeeHMC* h, int cycles, double* U) {
  HIP_CHECK(hipStreamSynchronize(h->stream));
  int done = 0;
  HIP_CHECK(hipMemcpy(&done, h->cycle.p, sizeof(int), hipMemcpyDeviceToHost));
  int n = std::max(0, std::min({ cycles, HMC_HIST, done }));
  if (n == 0 || !U) return n;
  std::vector<double> all((size_t)HMC_HIST * h->nrep);
  HIP_CHECK(hipMemcpy(all.data(), h->Uhist.p, sizeof(double) * all.size(), hipMemcpyDeviceToHost));
  for (int c = 0; c < n; c++) {
    int row = (done - n + c) % HMC_HIST;
    std::copy(all.begin() + (size_t)row * h->nrep, all.begin() + (size_t)(row + 1) * h->nrep, U + (size_t)c * h->nrep);
  }
  return n;
}

// Chosen variant {block size, cell division} and the autotuner's timings.
void EmDeeHMC_variant(EmDeeHMC* h, int* blockSize, int* cellDivision, double* tuneMs) {
  if (blockSize) *blockSize = VARIANTS[h->variant].bs;
  if (cellDivision) *cellDivision = VARIANTS[h->variant].div;
  if (tuneMs) for (int v = 0; v < NVARIANTS; v++) tuneMs[v] = h->tuneMs[v];
}

void EmDeeHMC_destroy(EmDeeHMC* h) {
  if (!h) return;
  (void)hipStreamSynchronize(h->stream);
  if (h->exec) (void)hipGraphExecDestroy(h->exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  h->desc.release(); h->order.release(); h->cycle.release(); h->U.release(); h->lastDH.release(); h->accepted.release();
  for (DBuf<double4>* b : { &h->R, &h->P, &h->F, &h->Rsave, &h->Fsave }) b->release();
  h->nlist.release(); h->nnb.release(); h->qbuild.release(); h->lstate.release(); h->Uhist.release();
  (void)hipStreamDestroy(h->stream);
  delete h;
}

// The generator of the momenta and acceptance uniforms, for hosts that
// replay a run (3 * n Gaussian momenta of unit mass at kT, then the uniform
// of the Metropolis test).
void EmDeeHMC_momenta(unsigned long long seed, int replica, int cycle, int n, double kT, double* P) {
  uint64_t stream = hmc_stream(replica, cycle);
  double sq = std::sqrt(kT);
  for (int i = 0; i < n; i++)
    for (int x = 0; x < 3; x++) {
      double u1 = hmc_uniform(seed, stream, 6ULL * i + 2 * x), u2 = hmc_uniform(seed, stream, 6ULL * i + 2 * x + 1);
      P[3 * i + x] = sq * (std::sqrt(-2.0 * std::log(u1)) * std::cos(TWO_PI_D * u2));
    }
}
double EmDeeHMC_uniform(unsigned long long seed, int replica, int cycle) {
  return hmc_uniform(seed, hmc_stream(replica, cycle), HMC_ACCEPT_COUNTER);
}

}  // extern "C"
