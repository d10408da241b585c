// device.hip -- HBM state and gfx950 kernels of the EmDee force engine.
//
// Layout in HBM (one system):
//   * every per-atom array is SoA and stored in *cell-sorted order*: slot s
//     holds original atom perm[s].  Coordinates R, momenta P, forces F (one
//     set per model layer), scaled coordinates Rs = R/L and the positions R0
//     at the latest list build are FP64 [3][Npad].
//   * the Verlet list is a FULL list in packets of four: entries 4p..4p+3 of
//     slot s are the 16 bytes at int offset 4 (p*Npad + s) (nb_at), so a lane
//     reads or writes four of its entries with one access and lanes of a
//     wavefront (consecutive slots) touch consecutive packets.  A full list
//     (every pair stored twice) lets one thread own one atom's force: no
//     atomics, deterministic sums.
//   * the list is rebuilt when the reference trigger fires
//     (neighbor_lists.f90:41-59, 175-195): cells of side >= (rc+skin)/3, a
//     counting sort of atoms by cell (stable: ascending original index within
//     a cell), then a scan of the 7 x 7 rows of cells around each atom's cell
//     (each row one contiguous slot range, rows outside the cutoff sphere
//     culled) with the reference's FP64 distance test on R/L.
//
// All launches go to one stream per system.  Host syncs happen only where
// the C ABI must return host-visible values (energies, virial, builds).
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "device_state.h"
#include "engine.h"
#include "pair_eval.h"

namespace emdee {

[[noreturn]] void fatal(const std::string& task, const std::string& msg) {
  fprintf(stderr, "Error in %s: %s.\n", task.c_str(), msg.c_str());
  fflush(stderr);
  exit(1);
}
void warning(const std::string& msg) { fprintf(stderr, "WARNING: %s.\n", msg.c_str()); }

// --------------------------------------------------------------- kernels

// Waits for the post with sequence number seq (k_reduce_post): polls the
// coherent host word, then falls back to a stream wait, which also reports a
// failed launch.
static void post_wait(Device* d, unsigned long long seq) {
  volatile unsigned long long* f = d->hseq;
  for (int k = 0; k < (1 << 20); k++) {
    if (*f >= seq) { std::atomic_thread_fence(std::memory_order_acquire); return; }
    __builtin_ia32_pause();
  }
  HIP_CHECK(hipStreamSynchronize(d->stream));
  if (*f < seq) fatal("device", "a posted result did not arrive");
  std::atomic_thread_fence(std::memory_order_acquire);
}

void dev_post_wait(Device* d, unsigned long long seq) { post_wait(d, seq); }
void dev_wait_posts(Device* d) { post_wait(d, d->seq); }

__global__ void k_iota(int* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = i;
}

// AoS (original order) -> SoA (sorted order)
__global__ void k_gather_aos(double* __restrict__ dst, const double* __restrict__ aos,
                             const int* __restrict__ perm, const int* __restrict__ mask, int N, int Npad) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  if (mask && !mask[s]) return;
  int o = perm[s];
  dst[s] = aos[3 * o];
  dst[Npad + s] = aos[3 * o + 1];
  dst[2 * Npad + s] = aos[3 * o + 2];
}
// SoA (sorted order) -> AoS (original order)
__global__ void k_scatter_aos(double* __restrict__ aos, const double* __restrict__ src,
                              const int* __restrict__ perm, int N, int Npad) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  int o = perm[s];
  aos[3 * o] = src[s];
  aos[3 * o + 1] = src[Npad + s];
  aos[3 * o + 2] = src[2 * Npad + s];
}
__global__ void k_add_aos(double* __restrict__ dst, const double* __restrict__ aos,
                          const int* __restrict__ perm, int N, int Npad) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  int o = perm[s];
  dst[s] += aos[3 * o];
  dst[Npad + s] += aos[3 * o + 1];
  dst[2 * Npad + s] += aos[3 * o + 2];
}
template <typename T>
__global__ void k_gather_scalar(T* __restrict__ dst, const T* __restrict__ src, const int* __restrict__ perm, int N) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < N) dst[s] = src[perm[s]];
}

// Rs = R / L, exactly as EmDeeCode.f90:1228 (IEEE division), packed with
// the atom type so that a neighbor costs one 32-byte gather.
__global__ void k_scale4(double4* __restrict__ Rs4, const double* __restrict__ R, const int* __restrict__ type,
                         double L, int N, int Npad) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < N) Rs4[s] = make_double4(R[s] / L, R[Npad + s] / L, R[2 * Npad + s] / L, (double)type[s]);
}

// ---- rebuild trigger, neighbor_lists.f90:41-59 ---------------------------
// With d_k = |R_k - R0_k|^2 in ORIGINAL atom order, the reference returns
// max + 2 sqrt(max*next) + next where next = d_1 if atom 1 holds the maximum
// and otherwise the maximum over atoms before the first one that attains it.
// Pass A: per block (max, smallest original index attaining it); its
// reduction gives the maximum and the first index k.  Pass C: per block max
// over atoms with original index < k.  Both reductions are deterministic.
// Scaled positions of a slot: Rs = R / L (EmDeeCode.f90:1228, IEEE division)
// with the type, and, when Q4 is given, the mixed form's fixed-point copy
// u = round(2^32 (s - floor s)) mod 2^32 and FP32 charge.
// Q4.w holds the type; a slot whose scaled position is not finite also
// carries Q4_NONFINITE, so that the fixed-point prefilter of the list build
// leaves its candidates to the FP64 test (which rejects them, as the
// reference's r^2 comparison does with NaN).
constexpr int Q4_NONFINITE = 1 << 30, Q4_TYPE = Q4_NONFINITE - 1;
struct Scaled {
  double4* Rs4;
  int4* Q4;
  float* qf;
  const int* type;
  const double* charge;
  const int* charged;
  double L;
};
__device__ __forceinline__ void scale_vals(const Scaled& sc, double x, double y, double z, int s) {
  double v[3] = { x / sc.L, y / sc.L, z / sc.L };
  int t = sc.type[s];
  if (sc.Rs4) sc.Rs4[s] = make_double4(v[0], v[1], v[2], (double)t);
  if (sc.Q4) {
    unsigned int u[3];
    for (int x = 0; x < 3; x++) {
      double f = v[x] - floor(v[x]);
      u[x] = (unsigned int)(unsigned long long)llrint(f * 4294967296.0);
    }
    bool fin = isfinite(v[0]) && isfinite(v[1]) && isfinite(v[2]);
    sc.Q4[s] = make_int4((int)u[0], (int)u[1], (int)u[2], fin ? t : (t | Q4_NONFINITE));
    if (sc.qf) sc.qf[s] = sc.charged[s] ? (float)sc.charge[s] : 0.0f;
  }
}
__device__ __forceinline__ void scale_slot(const Scaled& sc, const double* __restrict__ R, int s, int Npad) {
  scale_vals(sc, R[s], R[Npad + s], R[2 * Npad + s], s);
}

// d_k of slot s, |R - R0|^2 summed in this order with no contraction (pass
// C recomputes it rather than storing it: it runs only near the threshold)
__device__ __forceinline__ double disp_sq(const double* __restrict__ R, const double* __restrict__ R0, int s, int Npad) {
#pragma clang fp contract(off)
  double dx = R[s] - R0[s], dy = R[Npad + s] - R0[Npad + s], dz = R[2 * Npad + s] - R0[2 * Npad + s];
  return dx * dx + dy * dy + dz * dz;
}

// Pass A also writes the scaled positions of the current coordinates (sc),
// which the pair kernels of the same force evaluation read.
__global__ void k_disp_a(double* __restrict__ part, const double* __restrict__ R, const double* __restrict__ R0,
                         const int* __restrict__ perm, int N, int Npad, Scaled sc) {
  __shared__ double smax[BLOCK / 64];
  __shared__ int sidx[BLOCK / 64];
  int s = blockIdx.x * BLOCK + threadIdx.x;
  double v = -1.0;
  int idx = INT_MAX;
  if (s < N) {
    v = disp_sq(R, R0, s, Npad);
    idx = perm[s];
    scale_slot(sc, R, s, Npad);
  }
  for (int off = 32; off > 0; off >>= 1) {
    double ov = __shfl_down(v, off, 64);
    int oi = __shfl_down(idx, off, 64);
    if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
  }
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) { smax[wid] = v; sidx[wid] = idx; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < BLOCK / 64; w++)
      if (smax[w] > v || (smax[w] == v && sidx[w] < idx)) { v = smax[w]; idx = sidx[w]; }
    part[2 * blockIdx.x] = v;
    part[2 * blockIdx.x + 1] = (double)idx;
  }
}
// (max, first index) of a set of pairs: the larger value wins, ties go to the
// smaller index, so any combination order gives the same pair.
__device__ __forceinline__ void max_first(double& v, double& idx, double ov, double oi) {
  if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
}
// ... and posts {max, first index} to the host as well (as k_reduce_post)
__global__ void k_disp_a_reduce(double* __restrict__ trig, const double* __restrict__ part, int nb,
                                double* __restrict__ host, unsigned long long* __restrict__ hostSeq,
                                unsigned long long seq) {
  __shared__ double smax[BLOCK / 64], sidx[BLOCK / 64];
  double v = -1.0, idx = 2147483647.0;
#pragma unroll 8
  for (int b = threadIdx.x; b < nb; b += BLOCK) max_first(v, idx, part[2 * b], part[2 * b + 1]);
  for (int off = 32; off > 0; off >>= 1) max_first(v, idx, __shfl_down(v, off, 64), __shfl_down(idx, off, 64));
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) { smax[w] = v; sidx[w] = idx; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < BLOCK / 64; k++) max_first(v, idx, smax[k], sidx[k]);
    trig[0] = v;
    trig[1] = idx;
    host[0] = v;
    host[1] = idx;
    __threadfence_system();
    *(volatile unsigned long long*)hostSeq = seq;
  }
}
__global__ void k_disp_c(double* __restrict__ part, const double* __restrict__ R, const double* __restrict__ R0,
                         const int* __restrict__ perm, const double* __restrict__ trig, int N, int Npad) {
  __shared__ double smax[BLOCK / 64];
  int s = blockIdx.x * BLOCK + threadIdx.x;
  double k = trig[1];
  double v = 0.0;
  if (s < N && (double)perm[s] < k) v = disp_sq(R, R0, s, Npad);
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  if ((threadIdx.x & 63) == 0) smax[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < BLOCK / 64; w++) v = fmax(v, smax[w]);
    part[blockIdx.x] = v;
  }
}
__global__ void k_disp_c_reduce(double* __restrict__ trig, const double* __restrict__ part, int nb,
                                double* __restrict__ host, unsigned long long* __restrict__ hostSeq,
                                unsigned long long seq) {
  __shared__ double smax[BLOCK / 64];
  double v = 0.0;
#pragma unroll 8
  for (int b = threadIdx.x; b < nb; b += BLOCK) v = fmax(v, part[b]);
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  if ((threadIdx.x & 63) == 0) smax[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {   // posts {max, k, next} to the host (as k_reduce_post)
    for (int k = 1; k < BLOCK / 64; k++) v = fmax(v, smax[k]);
    trig[2] = v;
    host[0] = trig[0]; host[1] = trig[1]; host[2] = v;
    __threadfence_system();
    *(volatile unsigned long long*)hostSeq = seq;
  }
}

// ---- cell sort ------------------------------------------------------------
// Cell of slot s: icoord = int(M*(Rs - floor(Rs))) (neighbor_lists.f90:126);
// a value that rounds to M is the periodic image 0.
__device__ __forceinline__ int cell_coord(double s, int M) {
  int c = (int)(M * (s - floor(s)));
  return c >= M ? c - M : c;
}
// Cell of each slot from Rs = R / L (the division of k_scale4), counted per
// cell for the counting sort.
__global__ void k_cell_key(int* __restrict__ key, int* __restrict__ count, const double* __restrict__ R, double L,
                           int M, int N, int Npad) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  int ix = cell_coord(R[s] / L, M), iy = cell_coord(R[Npad + s] / L, M), iz = cell_coord(R[2 * Npad + s] / L, M);
  int c = ix + M * (iy + M * iz);
  key[s] = c;
  atomicAdd(&count[c], 1);
}
// Exclusive scan, three phases: per-block scan, scan of block sums, add.
__global__ void k_scan_block(int* __restrict__ out, int* __restrict__ bsum, const int* __restrict__ in, int n) {
  __shared__ int sh[BLOCK];
  int i = blockIdx.x * BLOCK + threadIdx.x;
  int v = i < n ? in[i] : 0;
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < BLOCK; off <<= 1) {
    int t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  if (i < n) out[i] = sh[threadIdx.x] - v;
  if (threadIdx.x == BLOCK - 1) bsum[blockIdx.x] = sh[BLOCK - 1];
}
// Exclusive scan of the nb block sums in one block of SCAN_T threads: each
// thread sums a contiguous run, the runs are scanned across the block, and
// each thread writes its run's prefixes.
constexpr int SCAN_T = 1024;
__global__ void __launch_bounds__(SCAN_T) k_scan_sums(int* bsum, int nb, int* total) {
  __shared__ int wsum[SCAN_T / 64];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int per = (nb + SCAN_T - 1) / SCAN_T, a = min(t * per, nb), b = min(a + per, nb);
  int sum = 0;
  for (int k = a; k < b; k++) sum += bsum[k];
  int incl = sum;
  for (int off = 1; off < 64; off <<= 1) {
    int y = __shfl_up(incl, off, 64);
    if (lane >= off) incl += y;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  if (t == 0) {
    int run = 0;
    for (int k = 0; k < SCAN_T / 64; k++) { int v = wsum[k]; wsum[k] = run; run += v; }
  }
  __syncthreads();
  int run = wsum[w] + incl - sum;
  for (int k = a; k < b; k++) { int v = bsum[k]; bsum[k] = run; run += v; }
  if (t == SCAN_T - 1 && total) *total = run;
}
__global__ void k_scan_add(int* __restrict__ out, const int* __restrict__ bsum, int n) {
  int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) out[i] += bsum[blockIdx.x];
}
__global__ void k_cell_scatter(int* __restrict__ order, int* __restrict__ fill, const int* __restrict__ key,
                               const int* __restrict__ start, int N) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  int c = key[s];
  order[start[c] + atomicAdd(&fill[c], 1)] = s;
}
// Deterministic order inside each cell: ascending original atom index.  Each
// entry of the scattered order counts the members of its cell with a smaller
// original index, which is its place in the cell (original indices are
// distinct); one thread per atom, so a few large cells do not serialise.
__global__ void k_cell_rank(int* __restrict__ order, const int* __restrict__ scat, const int* __restrict__ key,
                            const int* __restrict__ start, const int* __restrict__ perm, int N) {
  int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= N) return;
  const int v = scat[x], c = key[v], a = start[c], b = start[c + 1], pv = perm[v];
  int rank = 0;
  for (int y = a; y < b; y++) rank += perm[scat[y]] < pv;
  order[a + rank] = v;
}
// ncomp consecutive [Npad] components at once
// Every per-slot array gathered into the new slot order in one pass: slot s
// takes old slot order[s].  R goes to both R (via scratch) and R0 (the list's
// reference positions); the inverse permutation is written as perm is.
// F is permuted only when given (other layers' forces; the current layer's
// are recomputed after every list build).
struct PermuteAll {
  const int* order;
  const double *R, *P, *F, *invMass, *charge;
  double *Rn, *R0, *Pn, *Fn, *invMassn, *chargen;
  const int *type, *body, *perm, *freeFlag, *chargedFlag;
  int *typen, *bodyn, *permn, *freeFlagn, *chargedFlagn, *iperm;
  int nF;
};
__global__ void k_permute_all(PermuteAll a, int N, int Npad) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  const int o = a.order[s];
  for (int c = 0; c < 3; c++) {
    double r = a.R[(size_t)c * Npad + o];
    a.Rn[(size_t)c * Npad + s] = r;
    a.R0[(size_t)c * Npad + s] = r;
    a.Pn[(size_t)c * Npad + s] = a.P[(size_t)c * Npad + o];
  }
  if (a.F)
    for (int c = 0; c < a.nF; c++) a.Fn[(size_t)c * Npad + s] = a.F[(size_t)c * Npad + o];
  a.invMassn[s] = a.invMass[o];
  a.chargen[s] = a.charge[o];
  a.typen[s] = a.type[o];
  a.bodyn[s] = a.body[o];
  a.freeFlagn[s] = a.freeFlag[o];
  a.chargedFlagn[s] = a.chargedFlag[o];
  const int p = a.perm[o];
  a.permn[s] = p;
  a.iperm[p] = s;
}

// ---- mixed form: FP32 pair arithmetic, FP64 accumulation -----------------
// Positions are also kept as 32-bit fixed point of the scaled coordinate,
// u = round(2^32 (s - floor(s))) mod 2^32.  The wrapped integer difference
// (int)(u_i - u_j) is the minimum-image separation in units of L/2^32, so a
// neighbor costs one 16-byte gather {ux, uy, uz, type} and three integer
// subtractions, with no rounding-to-image step.  Converted to FP32 it carries
// a relative error below 1e-7.  Pairs whose FP32 r^2 falls within 2e-6 of the
// cutoff are decided by the FP64 test of the plain form on the FP64 scaled
// positions, so both forms keep exactly the same set of interacting pairs.
// The band is 2e-6 (FP32 rounding) plus the quantisation of the coordinates
// (half a unit each, at most 2^-32 of L per component of a separation):
// |r^2 - r_q^2| <= 2 sqrt(3) 2^-32 r + 3 2^-64, relative to Rc^2 below
// 4 2^-32 / (Rc/L) + 3 2^-64 / (Rc/L)^2 (band_rel).  The Verlet list build
// uses the same fixed-point test with the same band.
// Pair forces are summed in FP32 over groups of four neighbors, and each
// group is added to FP64 accumulators.
struct MixedCut {
  float K;          // (2^32 / L)^2: invR2 (length^-2) = K / r2 (fixed-point units)
  float lo, hi;     // FP32 band around the cutoff in fixed-point units^2
  double Rc2s;      // FP64 cutoff on scaled coordinates (the plain form's)
  double unit;      // L / 2^32
  double L;
};

// P16 (optional): the top 16 bits of each fixed-point coordinate, x and y
// packed in one word, z with bit 31 set when the position is not finite.
// key (optional): the cell of each slot, as k_cell_key computes it.
__global__ void k_scale4q(double4* __restrict__ Rs4, int4* __restrict__ Q4, float* __restrict__ qf,
                          const double* __restrict__ R, const int* __restrict__ type, const double* __restrict__ charge,
                          const int* __restrict__ charged, double L, int N, int Npad, uint2* __restrict__ P16,
                          int* __restrict__ key, int M) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  Scaled sc{ Rs4, Q4, qf, type, charge, charged, L };
  scale_slot(sc, R, s, Npad);
  if (key) {
    double4 r = Rs4[s];
    key[s] = cell_coord(r.x, M) + M * (cell_coord(r.y, M) + M * cell_coord(r.z, M));
  }
  if (P16) {
    int4 q = Q4[s];
    P16[s] = make_uint2(((unsigned int)q.x >> 16) | ((unsigned int)q.y & 0xffff0000u),
                        ((unsigned int)q.z >> 16) | ((q.w & Q4_NONFINITE) ? 0x80000000u : 0u));
  }
}

// relative half-width of the FP32 band around a cutoff Rc2s (scaled units^2)
inline double band_rel(double Rc2s) {
  const double q = 1.0 / 4294967296.0;
  return 2e-6 + 4.0 * q / std::sqrt(Rc2s) + 3.0 * q * q / Rc2s;
}

// The same test from the unscaled positions: Rs = R / L as k_scale4 forms it.
__device__ __noinline__ bool exact_inside_R(const double* __restrict__ R, int Npad, double L, int i, int j,
                                           double Rc2s) {
  double dx = R[i] / L - R[j] / L, dy = R[Npad + i] / L - R[Npad + j] / L, dz = R[2 * Npad + i] / L - R[2 * Npad + j] / L;
  dx = dx - rint(dx); dy = dy - rint(dy); dz = dz - rint(dz);
  double r2;
  {
#pragma clang fp contract(off)
    r2 = dx * dx + dy * dy + dz * dz;
  }
  return r2 < Rc2s;
}

__device__ __noinline__ bool exact_inside(const double4* __restrict__ Rs4, int i, int j, double Rc2s) {
  double4 a = Rs4[i], b = Rs4[j];
  double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  dx = dx - rint(dx); dy = dy - rint(dy); dz = dz - rint(dz);
  double r2;
  {
#pragma clang fp contract(off)
    r2 = dx * dx + dy * dy + dz * dz;
  }
  return r2 < Rc2s;
}

// ---- Verlet list ------------------------------------------------------------
constexpr int CELL_REACH = 3;
// Full list: for slot i, every j != i in the cells within +-CELL_REACH cells
// (cell edge >= (rc+skin)/CELL_REACH) with r^2 < (rc+skin)^2/L^2 computed as in
// neighbor_lists.f90:244 (FP64, minimum image by anint, no FMA contraction),
// that is not excluded, not in the same rigid body and whose types interact
// (neighbor_lists.f90:250-259).  The reference uses cells of edge rc+skin and
// the 27 around; thinner cells cut the candidates per atom (about 140 for 78
// neighbours at LJ liquid density, against 190 with edge (rc+skin)/2) while
// the exact test keeps the same pairs.  Slots are sorted by cell with x
// fastest, so for each (dy, dz) offset the cells x-R..x+R form one contiguous
// slot range (two when the row wraps); lanes of a wave (neighbouring slots)
// scan nearly the same ranges, which keeps the gathers in L1.  Candidates are loaded four at a
// time so that the loads of a lane overlap; they are still tested and
// appended in ascending slot order.
// SIMPLE: no rigid bodies, no exclusions and every type pair interacts, so
// the distance test alone decides (free atoms have distinct body ids); it is
// made on the 16-byte fixed-point positions, and only candidates within the
// FP32 band around the cutoff take the FP64 test on the scaled positions.
// S > 1: the S consecutive lanes of an atom take S consecutive candidates at
// a time; a ballot over the group places the accepted ones, so the list
// keeps the ascending order of the one-lane scan (for long lists: water).
// Pair test of the multi-lane scans, on the candidate's fixed-point position
// qj and body bj loaded by the caller: the fixed-point prefilter with the FP64
// test in the band, then (unless SIMPLE) the body, type and exclusion filters.
template <bool SIMPLE>
__device__ __forceinline__ bool list_accept(int j, int4 qj, int bj, int i, int4 qi, int ti, int bi,
                                            const double4* __restrict__ Rs4, const int* __restrict__ perm,
                                            const uint8_t* __restrict__ interact, const int* __restrict__ exItem,
                                            int e0, int e1, int ntypes, double xRc2, float lo, float hi) {
  float dx = (float)(int)((unsigned int)qi.x - (unsigned int)qj.x);
  float dy = (float)(int)((unsigned int)qi.y - (unsigned int)qj.y);
  float dz = (float)(int)((unsigned int)qi.z - (unsigned int)qj.z);
  float r2 = dx * dx + dy * dy + dz * dz;
  bool odd = ((qi.w | qj.w) & Q4_NONFINITE) != 0;
  bool in = r2 < lo && !odd;
  if (!in && (r2 < hi || odd)) in = exact_inside(Rs4, i, j, xRc2);
  bool ok = in & (j != i);
  if (SIMPLE) return ok;
  ok = ok & (bj != bi);
  if (interact) ok = ok && interact[(qj.w & Q4_TYPE) * ntypes + ti];   // null: all types interact
  if (ok && e1 > e0) {
    int oj = perm[j];
    for (int e = e0; e < e1; e++) if (exItem[e] == oj) return false;
  }
  return ok;
}
template <bool SIMPLE, int S>
__device__ __forceinline__ void scan_range(int a, int b, int i, double xi, double yi, double zi, int ti, int bi, int oi,
                                           const double4* __restrict__ Rs4, const int* __restrict__ body,
                                           const int* __restrict__ perm, const uint8_t* __restrict__ interact,
                                           const int* __restrict__ exItem, int e0, int e1, int ntypes, double xRc2,
                                           int cap, int Npad, int* __restrict__ nbr, int& cnt,
                                           const int4* __restrict__ Q4, int4 qi, float lo, float hi) {
  if (S > 1) {
    auto accept = [&](int j) -> bool {
      return list_accept<SIMPLE>(j, Q4[j], SIMPLE ? 0 : body[j], i, qi, ti, bi, Rs4, perm, interact, exItem, e0, e1,
                                 ntypes, xRc2, lo, hi);
    };
    const int lane = threadIdx.x & 63, sub = lane & (S - 1), gbase = lane & ~(S - 1);
    for (int j0 = a; j0 < b; j0 += S) {
      int j = j0 + sub;
      bool ok = j < b && accept(j);
      unsigned int g = (unsigned int)(__ballot(ok) >> gbase) & ((1u << S) - 1u);
      int pos = cnt + __popc(g & ((1u << sub) - 1u));
      if (ok && pos < cap) nbr[nb_at(pos, i, Npad)] = j;
      cnt += __popc(g);
    }
    return;
  }
  if (SIMPLE) {
    // rows are short (a few candidates), so every row is taken four
    // candidates at a time with validity masks; the rare band test is
    // hoisted out of the per-candidate path
    for (int j = a; j < b; j += 4) {
      int4 q[4];
#pragma unroll
      for (int u = 0; u < 4; u++) q[u] = Q4[min(j + u, b - 1)];
      bool in[4], band = false;
#pragma unroll
      for (int u = 0; u < 4; u++) {
        float dx = (float)(int)((unsigned int)qi.x - (unsigned int)q[u].x);
        float dy = (float)(int)((unsigned int)qi.y - (unsigned int)q[u].y);
        float dz = (float)(int)((unsigned int)qi.z - (unsigned int)q[u].z);
        float r2 = dx * dx + dy * dy + dz * dz;
        bool odd = ((qi.w | q[u].w) & Q4_NONFINITE) != 0;
        bool fast = r2 < lo && !odd;
        bool valid = j + u < b;
        in[u] = valid && fast && j + u != i;
        band = band || (valid && !fast && (r2 < hi || odd));
      }
      if (band) {
#pragma unroll
        for (int u = 0; u < 4; u++) {
          if (j + u < b && !in[u] && j + u != i) in[u] = exact_inside(Rs4, i, j + u, xRc2);
        }
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        if (in[u]) {
          if (cnt < cap) nbr[nb_at(cnt, i, Npad)] = j + u;
          cnt++;
        }
      }
    }
    return;
  }
  auto test = [&](int j, double4 rj) {
    double r2;
    {
#pragma clang fp contract(off)
      double d0 = xi - rj.x, d1 = yi - rj.y, d2 = zi - rj.z;
      d0 = d0 - rint(d0); d1 = d1 - rint(d1); d2 = d2 - rint(d2);
      r2 = d0 * d0 + d1 * d1 + d2 * d2;
    }
    if (r2 < xRc2 && j != i) {
      if (!SIMPLE) {
        if (body[j] == bi) return;
        if (interact && !interact[(int)rj.w * ntypes + ti]) return;
        if (e1 > e0) {
          int oj = perm[j];
          for (int e = e0; e < e1; e++) if (exItem[e] == oj) return;
        }
      }
      if (cnt < cap) nbr[nb_at(cnt, i, Npad)] = j;
      cnt++;
    }
  };
  int j = a;
  for (; j + 3 < b; j += 4) {
    double4 r0 = Rs4[j], r1 = Rs4[j + 1], r2 = Rs4[j + 2], r3 = Rs4[j + 3];
    test(j, r0); test(j + 1, r1); test(j + 2, r2); test(j + 3, r3);
  }
  for (; j < b; j++) test(j, Rs4[j]);
}
template <bool SIMPLE, int S>
__global__ void __launch_bounds__(BLOCK)
k_build_list(int* __restrict__ nbr, int* __restrict__ nnb, int* __restrict__ flags, double* __restrict__ part,
             const double4* __restrict__ Rs4, const int* __restrict__ key, const int* __restrict__ start,
             const int* __restrict__ body, const int* __restrict__ perm,
             const uint8_t* __restrict__ interact, const int* __restrict__ exOff, const int* __restrict__ exItem,
             int nexcl, int ntypes, int M, double xRc2, int cap, int N, int Npad, const int4* __restrict__ Q4,
             float lo, float hi) {
  int blk = xcd_block();
  int gt = blk * BLOCK + threadIdx.x;
  int i = S == 1 ? gt : gt / S, sub = S == 1 ? 0 : gt % S;
  int cnt = 0;
  if (i < N) {
    double4 ri = Rs4[i];
    int c = key[i];
    int cx = c % M, cy = (c / M) % M, cz = c / (M * M);
    int ti = (int)ri.w, bi = body[i], oi = perm[i];
    int4 qi = (SIMPLE || S > 1) ? Q4[i] : int4{ 0, 0, 0, 0 };
    int e0 = 0, e1 = 0;
    if (nexcl) { e0 = exOff[oi]; e1 = exOff[oi + 1]; }
    // Conservative culling: with f the fractional position of atom i inside
    // its cell, the row of cells at offset (dy, dz) is at least (gy, gz) cells
    // away, and inside that row only cells x with |x - (cx + fx)| <= e can
    // hold an atom closer than the cutoff rcell (cell units).  The exact FP64
    // test above still decides every scanned candidate; the margins absorb
    // the rounding of this estimate, so the pair set is unchanged.
    auto frac = [&](double v, int cc) {
      double f = M * (v - floor(v)) - cc;
      if (f > 1.5) f -= M;                  // the coordinate wrapped to cell 0
      return fmin(fmax(f, 0.0), 1.0);
    };
    // the culling arithmetic is FP32 with margins (1e-5 relative, 1e-4
    // cells) far above its rounding, so it can only keep extra cells
    const float fx = (float)frac(ri.x, cx), fy = (float)frac(ri.y, cy), fz = (float)frac(ri.z, cz);
    const float rcell2 = (float)(xRc2 * (double)M * (double)M * (1.0 + 1e-5) + 1e-5);
    // slot ranges of row r of the (dy, dz) offsets: cells xa..xb of the row,
    // split in two where the row wraps; empty when the row is culled
    struct Ranges { int a0, b0, a1, b1; };
    constexpr int W = 2 * CELL_REACH + 1;
    auto row = [&](int r) {
      Ranges q = { 0, 0, 0, 0 };
      int dz = r / W - CELL_REACH, dy = r % W - CELL_REACH;
      float gz = fmaxf(dz > 0 ? dz - fz : (dz < 0 ? fz - dz - 1 : 0.0f), 0.0f);
      float gy = fmaxf(dy > 0 ? dy - fy : (dy < 0 ? fy - dy - 1 : 0.0f), 0.0f);
      float rem = rcell2 - gy * gy - gz * gz;
      if (rem <= 0.0f) return q;
      float e = sqrtf(rem) + 1e-4f;
      int xa = cx + max((int)floorf(fx - e), -CELL_REACH), xb = cx + min((int)floorf(fx + e), CELL_REACH);
      int y = cy + dy; y = y < 0 ? y + M : (y >= M ? y - M : y);
      int z = cz + dz; z = z < 0 ? z + M : (z >= M ? z - M : z);
      int base = M * (y + M * z);
      int c0 = xa, c1 = xb, d0 = 0, d1 = -1;
      if (xa < 0 && xb >= 0) { c0 = xa + M; c1 = M - 1; d1 = xb; }
      else if (xa < 0) { c0 = xa + M; c1 = xb + M; }
      else if (xb >= M && xa < M) { c1 = M - 1; d1 = xb - M; }
      else if (xa >= M) { c0 = xa - M; c1 = xb - M; }
      q.a0 = start[base + c0]; q.b0 = start[base + c1 + 1];
      if (d1 >= d0) { q.a1 = start[base + d0]; q.b1 = start[base + d1 + 1]; }
      return q;
    };
    auto scan = [&](int a, int b) {
      scan_range<SIMPLE, S>(a, b, i, ri.x, ri.y, ri.z, ti, bi, oi, Rs4, body, perm, interact, exItem, e0, e1, ntypes, xRc2,
                         cap, Npad, nbr, cnt, Q4, qi, lo, hi);
    };
    if constexpr (S == 16) {
      // Streamed: the atom's 16 lanes form its rows 16 at a time (lane sub
      // takes rows sub, sub + 16, ...) and keep the non-empty ranges, in scan
      // order, in the group's slice of LDS.  Each step then takes the next 16
      // candidates of the concatenated ranges, so no step ends half-used at a
      // range's end and no lane repeats another's row arithmetic.  The
      // candidates, their order and the ballot placement are those of the
      // row-by-row scan, so the list is the same.
      constexpr int NROW = W * W, SEGS = 2 * NROW;
      __shared__ int2 segs[BLOCK / S][SEGS];
      int2* sg = segs[threadIdx.x / S];
      const int lane = threadIdx.x & 63, gbase = lane & ~(S - 1);
      const unsigned lower = (1u << sub) - 1u, gmask = (1u << S) - 1u;
      int nseg = 0, len = 0;
      for (int r0 = 0; r0 < NROW; r0 += S) {
        int r = r0 + sub;
        Ranges q = r < NROW ? row(r) : Ranges{ 0, 0, 0, 0 };
        int l0 = q.b0 - q.a0, l1 = q.b1 - q.a1;
        len += l0 + l1;
        unsigned g1 = (unsigned)(__ballot(l0 > 0 || l1 > 0) >> gbase) & gmask;
        unsigned g2 = (unsigned)(__ballot(l0 > 0 && l1 > 0) >> gbase) & gmask;
        int k = nseg + __popc(g1 & lower) + __popc(g2 & lower);
        if (l0 > 0) sg[k++] = make_int2(q.a0, l0);
        if (l1 > 0) sg[k] = make_int2(q.a1, l1);
        nseg += __popc(g1) + __popc(g2);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      // candidates of the atom, hence its number of steps
      for (int o = S / 2; o > 0; o >>= 1) len += __shfl_xor(len, o, S);
      const int steps = (len + S - 1) / S;
      // lane position: offset off inside range s, which starts at slot ca and
      // holds cl candidates; past the last range cl is INT_MAX
      int s = 0, off = sub, ca = 0, cl = INT_MAX;
      if (nseg > 0) { int2 t = sg[0]; ca = t.x; cl = t.y; }
      auto settle = [&]() {
        while (off >= cl) {
          off -= cl;
          s++;
          const int2 t = s < nseg ? sg[s] : make_int2(0, INT_MAX);
          ca = t.x; cl = t.y;
        }
      };
      settle();
      // one step ahead: the next step's position and body are loaded before
      // this step is tested (a lane without a candidate loads its own atom's)
      // (32-bit offsets from the array bases: slots are below 2^21 here)
      auto q4At = [&](int k) { return *(const int4*)((const char*)Q4 + (unsigned)k * 16u); };
      auto bodyAt = [&](int k) { return *(const int*)((const char*)body + (unsigned)k * 4u); };
      bool has = s < nseg;
      int j = has ? ca + off : i;
      int4 qj = q4At(j);
      int bj = SIMPLE ? 0 : bodyAt(j);
      for (int step = 0; step < steps; step++) {
        off += S;
        settle();
        const bool hasN = s < nseg;
        const int jN = hasN ? ca + off : i;
        const int4 qN = q4At(jN);
        const int bN = SIMPLE ? 0 : bodyAt(jN);
        bool ok = has && list_accept<SIMPLE>(j, qj, bj, i, qi, ti, bi, Rs4, perm, interact, exItem, e0, e1, ntypes,
                                             xRc2, lo, hi);
        unsigned g = (unsigned)(__ballot(ok) >> gbase) & gmask;
        int pos = cnt + __popc(g & lower);
        if (ok && pos < cap) nbr[nb_at(pos, i, Npad)] = j;
        cnt += __popc(g);
        has = hasN; j = jN; qj = qN; bj = bN;
      }
    } else {
      // the cell starts of the next row are loaded before the current row is
      // scanned, so their latency overlaps the scan
      Ranges cur = row(0);
      for (int r = 0; r < W * W; r++) {
        Ranges nxt = r + 1 < W * W ? row(r + 1) : Ranges{ 0, 0, 0, 0 };
        scan(cur.a0, cur.b0);
        scan(cur.a1, cur.b1);
        cur = nxt;
      }
    }
    if (sub == 0) {
      nnb[i] = cnt;
      // pad the list to a multiple of 4 with the atom itself (zero separation):
      // the mixed-form kernel reads whole groups of four without a bound test
      if (cnt <= cap)
        for (int k = cnt; k < ((cnt + 3) & ~3); k++) nbr[nb_at(k, i, Npad)] = i;
      if (cnt > cap) atomicMax(&flags[0], cnt);
    }
  }
  double v[1] = { sub == 0 ? (double)cnt : 0.0 };
  block_sum<1>(v);
  if (threadIdx.x == 0) part[blk] = v[0];
}

// The same pair set for free atoms with every type pair interacting
// (SIMPLE), one lane per atom, built for latency rather than by rows:
// * A wave that walks the 7x7 rows one by one pays, in every row, for its
//   lane with the most candidates there.  Here each lane loads the slot
//   bounds of the seven rows of a z plane at once, keeps the non-empty
//   segments in its own slice of LDS and streams its candidates four at a
//   time, so a wave pays for its busiest lane once per plane.
// * Prefilter: the 16-bit fixed-point coordinates of P16 give the squared
//   distance r2 (units 2^-32 box^2) in four instructions, wrapped
//   differences being the minimum image.  With e = 2^-16 + 2^-32 the largest
//   error of a coordinate difference, |r2 - r^2| <= 2 sqrt(3) e r + 3 e^2, so
//   r2 < lo (r2 >= hi) decides "inside" ("outside") for every candidate whose
//   exact r^2 is not within the relative band b of the cutoff.  Cells have
//   edge >= 1/12 of the box (M >= 12), so |differences| <= 4 cells < 2^15
//   and r2 < 2^31.
// * The candidates in the band (about 0.5 per atom) are queued in four
//   registers and decided after the stream by the FP64 test on the scaled positions, the
//   test of k_build_list, so the pair set is the same; they are appended
//   after the others.  The stream itself then loads only P16 and keeps two
//   groups' loads in flight, which its stores do not hold up (a wait for a
//   load also waits for every store issued before it).
constexpr int SEG_MAX = 2 * (2 * CELL_REACH + 1) + 1;   // segments of one plane (the own row split)
constexpr int BANDQ = 4;                             // queued band candidates per lane
typedef short short2v __attribute__((ext_vector_type(2)));
__global__ void __launch_bounds__(BLOCK)
k_build_list_seg(int* __restrict__ nbr, int* __restrict__ nnb, int* __restrict__ flags, double* __restrict__ part,
                 const double4* __restrict__ Rs4, const int* __restrict__ key, const int* __restrict__ start, int M,
                 double xRc2, int cap, int N, int Npad, const uint2* __restrict__ P16, int lo, int hi) {
  __shared__ int2 seg[SEG_MAX * BLOCK];
  __shared__ int ring[8 * BLOCK];
  int blk = xcd_block();
  int i = blk * BLOCK + threadIdx.x;
  int cnt = 0;
  if (i < N) {
    const double4 ri = Rs4[i];
    const int c = key[i];
    const int cx = c % M, cy = (c / M) % M, cz = c / (M * M);
    const uint2 pi = P16[i];
    const bool oddi = (int)pi.y < 0;
    auto frac = [&](double v, int cc) {
      double f = M * (v - floor(v)) - cc;
      if (f > 1.5) f -= M;                  // the coordinate wrapped to cell 0
      return fmin(fmax(f, 0.0), 1.0);
    };
    // conservative culling in FP32 with margins, as in k_build_list
    const float fx = (float)frac(ri.x, cx), fy = (float)frac(ri.y, cy), fz = (float)frac(ri.z, cz);
    const float rcell2 = (float)(xRc2 * (double)M * (double)M * (1.0 + 1e-5) + 1e-5);
    // gap, in cells, between the atom and the cells at offset d
    auto gap = [](int d, float f) { return fmaxf(fmaxf((float)d - f, f - (float)d - 1.0f), 0.0f); };
    constexpr int W = 2 * CELL_REACH + 1;
    float gy2[W];
    int yrow[W];
#pragma unroll
    for (int t = 0; t < W; t++) {
      float g = gap(t - CELL_REACH, fy);
      gy2[t] = g * g;
      int y = cy + t - CELL_REACH;
      yrow[t] = M * (y < 0 ? y + M : (y >= M ? y - M : y));
    }
    int2* __restrict__ my = seg + threadIdx.x;
    // queued band candidates, in four named registers (an array would be
    // indexed by nq and kept in scratch memory)
    int bq0 = 0, bq1 = 0, bq2 = 0, bq3 = 0, nq = 0;
    // Entries go to a lane-private ring of eight in LDS (entry c at slot
    // c % 8); a packet of four is stored with one 16-byte store once it is
    // complete.  Flushes follow each group of at most four appends, so at
    // most one packet completes between two flushes and the other half of
    // the ring is never overwritten before it is stored.
    int* __restrict__ rg = ring + threadIdx.x;   // ring slot r at rg[r * BLOCK]
    int4v* __restrict__ pk = reinterpret_cast<int4v*>(nbr) + i;   // packet p at pk[p * Npad]
    const int capP = cap >> 2;
    auto append = [&](int j) __attribute__((always_inline)) {
      rg[(cnt & 7) * BLOCK] = j;
      cnt++;
    };
    auto flush = [&](int before) __attribute__((always_inline)) {
      const int p = before >> 2;
      if ((cnt >> 2) > p && p < capP) {
        const int r = (p & 1) * 4;
        pk[(size_t)p * Npad] = (int4v){ rg[r * BLOCK], rg[(r + 1) * BLOCK], rg[(r + 2) * BLOCK], rg[(r + 3) * BLOCK] };
      }
    };
    auto exact = [&](int j) __attribute__((always_inline)) {
      double4 rj = Rs4[j];
      double r2;
      {
#pragma clang fp contract(off)
        double d0 = ri.x - rj.x, d1 = ri.y - rj.y, d2 = ri.z - rj.z;
        d0 = d0 - rint(d0); d1 = d1 - rint(d1); d2 = d2 - rint(d2);
        r2 = d0 * d0 + d1 * d1 + d2 * d2;
      }
      return r2 < xRc2;
    };
    for (int dz = -CELL_REACH; dz <= CELL_REACH; dz++) {
      float gz = gap(dz, fz);
      float remz = rcell2 - gz * gz;
      // slot bounds of the plane's rows, all loaded before any is used:
      // row t gives cells c0..e1 and, where it wraps, 0..d1 after them (the
      // order of k_build_list); a culled row gives empty ranges.  xa >= -3
      // and xa <= cx < M; xb >= cx >= 0 and xb <= cx + 3.
      int z = cz + dz; z = z < 0 ? z + M : (z >= M ? z - M : z);
      const unsigned int plane = (unsigned int)(M * M * z);
      int sa[SEG_MAX], sb[SEG_MAX];
#pragma unroll
      for (int t = 0; t < W; t++) {
        float rem = remz - gy2[t];
        bool on = rem > 0.0f;
        float e = __builtin_amdgcn_sqrtf(fmaxf(rem, 0.0f)) + 1e-4f;
        int xa = cx + max((int)floorf(fx - e), -CELL_REACH), xb = cx + min((int)floorf(fx + e), CELL_REACH);
        bool wlo = xa < 0, two = on && (wlo || xb >= M);
        unsigned int row = plane + (unsigned int)yrow[t];
        unsigned int i0 = row + (unsigned int)(on ? (wlo ? xa + M : xa) : 0);
        unsigned int i1 = row + (unsigned int)(on ? (wlo ? M : min(xb + 1, M)) : 0);
        unsigned int i3 = row + (unsigned int)(two ? (wlo ? xb + 1 : xb - M + 1) : 0);
        // slots of the row, in order; a culled row loads nothing, a row that
        // does not wrap loads only its first range
        int s0 = 0, s1 = 0, w0 = 0, w1 = 0;
        if (on) { s0 = start[i0]; s1 = start[i1]; }
        if (two) { w0 = start[row]; w1 = start[i3]; }
        if (t < CELL_REACH) {
          sa[2 * t] = s0; sb[2 * t] = s1; sa[2 * t + 1] = w0; sb[2 * t + 1] = w1;
        } else if (t > CELL_REACH) {
          sa[2 * t + 1] = s0; sb[2 * t + 1] = s1; sa[2 * t + 2] = w0; sb[2 * t + 2] = w1;
        } else {
          // the atom's own row, split at the atom: its cell is in the wrapped
          // part when the row wraps below 0, else in the first.  A slot
          // outside [p0, p1) leaves [p0, min(i, p1)) whole and the rest empty,
          // or the reverse, so the planes dz != 0 are split harmlessly.
          int p0 = wlo ? w0 : s0, p1 = wlo ? w1 : s1;
          int h0 = p0, h1 = min(i, p1), g0 = max(i + 1, p0), g1 = p1;
          sa[2 * t] = wlo ? s0 : h0;     sb[2 * t] = wlo ? s1 : h1;
          sa[2 * t + 1] = wlo ? h0 : g0; sb[2 * t + 1] = wlo ? h1 : g1;
          sa[2 * t + 2] = wlo ? g0 : w0; sb[2 * t + 2] = wlo ? g1 : w1;
        }
      }
      // the non-empty ones, in order, into the lane's LDS slice, so that the
      // stream below steps to the next segment with a single test
      int nseg = 0;
#pragma unroll
      for (int t = 0; t < SEG_MAX; t++) {
        my[BLOCK * nseg] = make_int2(sa[t], sb[t]);
        nseg += sa[t] < sb[t] ? 1 : 0;
      }
      if (oddi) nseg = 0;                   // a non-finite atom has no neighbours
      // stream the plane's segments (a lane-private LDS slice: no barrier);
      // group g is tested after the loads of groups g+1 and g+2 are issued
      int k = 0, a = 0, b = 0;
      if (nseg > 0) { int2 t = my[0]; a = t.x; b = t.y; }
      auto advance = [&]() __attribute__((always_inline)) {
        a += 4;
        if (a >= b && ++k < nseg) { int2 t = my[BLOCK * k]; a = t.x; b = t.y; }
      };
      // loads of the group at the current position; P16 has 16 readable
      // slots past the last atom, so a group (or the three past a plane's
      // end) never reads out of bounds
      auto fetch = [&](uint2 (&q)[4], int& qa, int& qb) __attribute__((always_inline)) {
        const bool more = k < nseg;
        qa = more ? a : 0;
        qb = more ? b : 0;
        const uint2* __restrict__ p = P16 + (unsigned int)a;
#pragma unroll
        for (int u = 0; u < 4; u++) q[u] = p[u];
        advance();
      };
      auto test = [&](const uint2 (&q)[4], int qa, int qb) __attribute__((always_inline)) {
        bool in[4], chk[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          short2v dxy = __builtin_bit_cast(short2v, pi.x) - __builtin_bit_cast(short2v, q[u].x);
          int dzs = (int)(short)(pi.y - q[u].y);
          int r2 = __builtin_amdgcn_sdot2(dxy, dxy, dzs * dzs, true);
          bool odd = (int)q[u].y < 0;
          bool ok = qa + u < qb;
          in[u] = ok && r2 < lo && !odd;
          chk[u] = ok && !in[u] && (r2 < hi || odd);
        }
        // one append site: increments of the count and of the queue length
        // on exclusive paths would otherwise be merged into a store through
        // a pointer, which keeps both in scratch memory
        const int before = cnt;
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int j = qa + u;
          bool take = in[u];
          if (chk[u]) {
            if (nq < BANDQ) {
              bq0 = nq == 0 ? j : bq0;
              bq1 = nq == 1 ? j : bq1;
              bq2 = nq == 2 ? j : bq2;
              bq3 = nq == 3 ? j : bq3;
            } else {
              take = exact(j);                // queue full (rare): decided now
            }
            nq = min(nq + 1, BANDQ + 1);
          }
          if (take) append(j);
        }
        flush(before);
      };
      uint2 q0[4], q1[4], q2[4];
      int a0, b0, a1, b1, a2, b2;
      fetch(q0, a0, b0);
      fetch(q1, a1, b1);
      // three register sets in turn: no copy waits for a load
      while (a0 < b0) {
        fetch(q2, a2, b2);
        test(q0, a0, b0);
        if (a1 >= b1) break;
        fetch(q0, a0, b0);
        test(q1, a1, b1);
        if (a2 >= b2) break;
        fetch(q1, a1, b1);
        test(q2, a2, b2);
      }
    }
    // the band candidates, by the FP64 test
    bool t0 = nq > 0 && exact(bq0), t1 = nq > 1 && exact(bq1), t2 = nq > 2 && exact(bq2), t3 = nq > 3 && exact(bq3);
    {
      const int before = cnt;
      if (t0) append(bq0);
      if (t1) append(bq1);
      if (t2) append(bq2);
      if (t3) append(bq3);
      flush(before);
    }
    nnb[i] = cnt;
    // the last packet, padded with the atom itself (zero separation): the
    // mixed-form kernel reads whole groups of four without a bound test
    if (cnt & 3) {
      const int n = cnt;
      while (cnt & 3) append(i);
      flush(n);
    }
    if (cnt > cap) atomicMax(&flags[0], cnt);
  }
  double v[1] = { (double)cnt };
  block_sum<1>(v);
  if (threadIdx.x == 0) part[blk] = v[0];
}

// integer band of k_build_list_seg for the squared list cutoff xRc2 (scaled
// units^2): {lo, hi} in units of 2^-32 box^2
static void band16(double xRc2, int* lo, int* hi) {
  const double e = 1.0 / 65536.0 + 1.0 / 4294967296.0;
  double r = std::sqrt(xRc2);
  double b = 2.0 * std::sqrt(3.0) * e / r + 3.0 * e * e / xRc2 + 1e-6;
  double c = xRc2 * 4294967296.0;
  *lo = (int)std::floor(c * (1.0 - b));
  *hi = (int)std::ceil(c * (1.0 + b));
}

// ---- pair forces ----------------------------------------------------------
// One thread per atom over its full list; the per-pair body is
// compute.f90:46-92.  Energies and virials of a full list count every pair
// twice and are halved at the end (exact).
//
// Minimum image: the reference subtracts anint(d) (half away from zero); rint
// (half to even, one v_rndne_f64) differs only at exact ties |d| = k + 1/2,
// where both give |d - n| = 1/2 -- the same r^2 -- and the pair is then at
// least L/2 apart, i.e. beyond any cutoff, so no force differs.

__device__ __forceinline__ double fast_rcp(double x) {
  // v_rcp_f64 + two Newton steps: within 1 ulp of 1/x
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  e = fma(-x, y, 1.0);
  return fma(y, e, y);
}

struct PairAcc { double fx, fy, fz, Ep, Ec, Wp, Wc; };

// Type-mixed systems of up to PAIR_LDS_TYPES types keep the pair-model table
// in LDS for the FP64 kernels, so that the record of every list entry is an
// LDS read rather than a dozen vector memory loads (98k-atom water: 0.63 ->
// 0.55 ms; the mixed-form kernel, whose loads are fewer, measured no gain).
// Every thread of the block calls this (it holds a barrier); the result is
// the same for the whole block.
constexpr int PAIR_LDS_TYPES = 8;
__device__ __forceinline__ bool stage_pairs(DevPair* spair, const DevPair* __restrict__ pairs, int ntypes) {
  if (ntypes > PAIR_LDS_TYPES) return false;
  for (int t = threadIdx.x; t < ntypes * ntypes; t += blockDim.x) spair[t] = pairs[t];
  __syncthreads();
  return true;
}
// |q| above which an atom is charged (api.cpp sets the flags the same way)
constexpr double Q_EPS = 2.220446049250313e-16;

// LJ: the model is pair_lj_cut with no modifier or an energy shift, and no
// pair is charged -- the single-model systems of the LJ workloads, compiled
// without the generic dispatch so that the kernel needs fewer registers.
template <typename T, bool ENERGY, bool LJ>
__device__ __forceinline__ void pair_body(const DevPair& p, const DevCoul& cm, bool ijcharged, double qiqj,
                                          double dx, double dy, double dz, double r2, double invL2, PairAcc& a) {
  T invR2;
  if (std::is_same<T, float>::value) invR2 = T(invL2) * rcp<T>(T(r2));
  else invR2 = T(invL2 * fast_rcp(r2));
  T E, W;
  if (LJ) {
    T sr2 = T(p.c2) * invR2, sr6 = sr2 * sr2 * sr2, sr12 = sr6 * sr6;
    W = T(p.c1) * (sr12 + sr12 - sr6);
    if (ENERGY) a.Ep += (double)(T(p.c0) * (sr12 - sr6) + T(p.eshift));
    a.Wp += (double)W;
    T g = W * invR2;
    a.fx += (double)(g * T(dx));
    a.fy += (double)(g * T(dy));
    a.fz += (double)(g * T(dz));
    return;
  }
  pair_model<T>(p, invR2, E, W);
  if (p.modifier != MOD_NONE) {
    T invR = (p.modifier == MOD_SHIFTED_FORCE || p.modifier == MOD_SMOOTHED || p.modifier == MOD_SHIFTED_SMOOTHED)
                 ? T(sqrt(invR2)) : T(0);
    apply_modifier<T>(p.modifier, T(p.eshift), T(p.fshift), T(p.factor), T(p.Rm2fac), invR, invR2, E, W);
  }
  if (ENERGY) a.Ep += (double)E;
  a.Wp += (double)W;
  T Wsum = W;
  if (ijcharged) {
    T invR = T(sqrt(invR2));
    coul_model<T, ENERGY>(cm, invR, invR2, E, W);
    apply_modifier<T>(cm.modifier, T(cm.eshift), T(cm.fshift), T(cm.factor), T(cm.Rm2fac), invR, invR2, E, W);
    if (ENERGY) a.Ec += qiqj * (double)E;
    T Wq = T(qiqj) * W;
    a.Wc += (double)Wq;
    Wsum = Wsum + Wq;
  }
  T g = Wsum * invR2;
  a.fx += (double)(g * T(dx));
  a.fy += (double)(g * T(dy));
  a.fz += (double)(g * T(dz));
}

// pair_body for a layer of lj_cut/none pair models (FP64): c = {c0, c1, c2,
// eshift} of the type pair (zeros for pair none; eshift zero without the
// shifted modifier), the same terms in the same order.
template <bool ENERGY>
__device__ __forceinline__ void pair_body_ljc(const double4& c, const DevCoul& cm, bool jc, double qiqj, double dx,
                                              double dy, double dz, double r2, double invL2, PairAcc& a) {
  double invR2 = invL2 * fast_rcp(r2);
  double sr2 = c.z * invR2, sr6 = sr2 * sr2 * sr2, sr12 = sr6 * sr6;
  double E = c.x * (sr12 - sr6), W = c.y * (sr12 + sr12 - sr6);
  E = E + c.w;
  if (ENERGY) a.Ep += E;
  a.Wp += W;
  double Wsum = W;
  if (jc) {
    double invR = sqrt(invR2);
    coul_model<double, ENERGY>(cm, invR, invR2, E, W);
    apply_modifier<double>(cm.modifier, cm.eshift, cm.fshift, cm.factor, cm.Rm2fac, invR, invR2, E, W);
    if (ENERGY) a.Ec += qiqj * E;
    double Wq = qiqj * W;
    a.Wc += Wq;
    Wsum = Wsum + Wq;
  }
  double g = Wsum * invR2;
  a.fx += g * dx;
  a.fy += g * dy;
  a.fz += g * dz;
}

// UNIFORM: a single atom type, so the model record is loaded once.
// TPA threads per atom (consecutive lanes) split an atom's list for small
// systems that would not fill the chip with one thread per atom; their
// partial forces are combined by a fixed butterfly (deterministic).
template <int TPA>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int off = TPA / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// LJ: 1 for the one-type LJ path; 2 for layers of lj_cut/none pair models
// with any Coulomb model, whose compact FP64 records (ljt: c0, c1, c2,
// eshift; ljk: kCoul or 0) are staged in LDS; 0 for every other layer.
template <bool ENERGY, bool MIXED, bool UNIFORM, int LJ, int TPA>
__global__ void __launch_bounds__(BLOCK)
k_pairs(double* __restrict__ F, double* __restrict__ partials,
        const double4* __restrict__ Rs4, const int* __restrict__ nbr, const int* __restrict__ nnb,
        const double* __restrict__ charge, const int* __restrict__ charged,
        const DevPair* __restrict__ pairs, const DevCoul* __restrict__ coulp,
        int ntypes, double L, double invL2, double Rc2, int N, int Npad, double* __restrict__ eatom,
        const double4* __restrict__ ljt, const double* __restrict__ ljk) {
  using T = typename std::conditional<MIXED, float, double>::type;
  int blk = xcd_block();
  int gt = blk * BLOCK + threadIdx.x;
  int i = TPA == 1 ? gt : gt / TPA, sub = TPA == 1 ? 0 : gt % TPA;
  PairAcc a = { 0, 0, 0, 0, 0, 0, 0 };
  __shared__ DevPair spair[PAIR_LDS_TYPES * PAIR_LDS_TYPES];
  __shared__ double4 sljt[PAIR_LDS_TYPES * PAIR_LDS_TYPES];
  __shared__ double sljk[PAIR_LDS_TYPES * PAIR_LDS_TYPES];
  bool lds = false;
  if (LJ == 2) {
    lds = ntypes <= PAIR_LDS_TYPES;
    if (lds) {
      for (int t = threadIdx.x; t < ntypes * ntypes; t += blockDim.x) { sljt[t] = ljt[t]; sljk[t] = ljk[t]; }
      __syncthreads();
    }
  } else if (!UNIFORM) {
    lds = stage_pairs(spair, pairs, ntypes);
  }
  if (i < N) {
    const DevCoul cm = *coulp;
    const double4 ri = Rs4[i];
    int ti = (int)ri.w;
    double qi = charge[i];
    bool ich = charged[i] != 0;
    const DevPair* prow = pairs + ti * ntypes;
    const DevPair* srow = spair + ti * ntypes;
    DevPair pu;
    if (UNIFORM) pu = pairs[0];   // uniform address: scalar loads
    const int n = nnb[i];
    // qjp: the neighbour's charge when the caller loaded it with the position
    // (LJ == 2 on the multi-lane path), else read here
    auto one = [&](int j, const double4& rj, double qjp = 0.0, bool haveQ = false) {
      double dx = ri.x - rj.x, dy = ri.y - rj.y, dz = ri.z - rj.z;
      dx = dx - rint(dx); dy = dy - rint(dy); dz = dz - rint(dz);
      double r2;
      {
#pragma clang fp contract(off)
        r2 = dx * dx + dy * dy + dz * dz;
      }
      if (r2 < Rc2) {
        if (LJ == 1) {
          pair_body<T, ENERGY, true>(pu, cm, false, 0.0, dx, dy, dz, r2, invL2, a);
        } else if (LJ == 2) {
          const int t = ti * ntypes + (int)rj.w;
          double4 c;
          double kc;
          if (lds) { c = sljt[t]; kc = sljk[t]; } else { c = ljt[t]; kc = ljk[t]; }
          double qj = haveQ ? qjp : charge[j];
          bool jc = ich && kc != 0.0 && fabs(qj) > Q_EPS;
          double qiqj = jc ? kc * qi * qj : 0.0;
          pair_body_ljc<ENERGY>(c, cm, jc, qiqj, dx, dy, dz, r2, invL2, a);
        } else if (UNIFORM) {
          bool jc = ich && pu.coulomb && charged[j];
          double qiqj = jc ? pu.kCoul * qi * charge[j] : 0.0;
          pair_body<T, ENERGY, false>(pu, cm, jc, qiqj, dx, dy, dz, r2, invL2, a);
        } else {
          DevPair p;
          if (lds) p = srow[(int)rj.w]; else p = prow[(int)rj.w];
          // charged means |q| > EPS (api.cpp): the charge alone decides, one load
          double qj = charge[j];
          bool jc = ich && p.coulomb && fabs(qj) > Q_EPS;
          double qiqj = jc ? p.kCoul * qi * qj : 0.0;
          pair_body<T, ENERGY, false>(p, cm, jc, qiqj, dx, dy, dz, r2, invL2, a);
        }
      }
    };
    if (TPA == 1) {
      int k = 0;
      for (; k + 4 <= n; k += 4) {
        const int4v jv = nb_packet(nbr, k >> 2, i, Npad);
        const int j0 = jv.x, j1 = jv.y, j2 = jv.z, j3 = jv.w;
        double4 r0 = Rs4[j0], r1 = Rs4[j1], r2 = Rs4[j2], r3 = Rs4[j3];
        one(j0, r0); one(j1, r1); one(j2, r2); one(j3, r3);
      }
      for (; k < n; k++) {
        int j = nbr[nb_at(k, i, Npad)];
        one(j, Rs4[j]);
      }
    } else {
      int k = sub;
      for (; k + 2 * TPA < n; k += 3 * TPA) {
        int j0 = nbr[nb_at(k, i, Npad)], j1 = nbr[nb_at(k + TPA, i, Npad)], j2 = nbr[nb_at(k + 2 * TPA, i, Npad)];
        double4 r0 = Rs4[j0], r1 = Rs4[j1], r2 = Rs4[j2];
        if (LJ == 2) {
          const double q0 = charge[j0], q1 = charge[j1], q2 = charge[j2];
          one(j0, r0, q0, true); one(j1, r1, q1, true); one(j2, r2, q2, true);
        } else {
          one(j0, r0); one(j1, r1); one(j2, r2);
        }
      }
      for (; k < n; k += TPA) { int j = nbr[nb_at(k, i, Npad)]; one(j, Rs4[j]); }
    }
  }
  if (TPA > 1) {
    a.fx = group_sum<TPA>(a.fx);
    a.fy = group_sum<TPA>(a.fy);
    a.fz = group_sum<TPA>(a.fz);
  }
  if (i < N && sub == 0) {
    F[i] = L * a.fx;
    F[Npad + i] = L * a.fy;
    F[2 * Npad + i] = L * a.fz;
  }
  if (ENERGY && eatom) {   // atom energies: half of every pair term of the atom
    double e = a.Ep + a.Ec;
    if (TPA > 1) e = group_sum<TPA>(e);
    if (i < N && sub == 0) eatom[i] = 0.5 * e;
  }
  double v[4] = { a.Ep, a.Ec, a.Wp, a.Wc };
  block_sum<4>(v);
  if (threadIdx.x == 0) {
    partials[4 * blk + 0] = 0.5 * v[0];
    partials[4 * blk + 1] = 0.5 * v[1];
    partials[4 * blk + 2] = 0.5 * v[2];
    partials[4 * blk + 3] = 0.5 * v[3];
  }
}

// ---- mixed form pair kernels (fixed-point helpers are above the list build) --

struct MixedAcc { float fx, fy, fz, Ep, Ec, Wp, Wc; };

template <bool ENERGY, bool LJ>
__device__ __forceinline__ void pair_body_mx(const DevPair& p, const DevCoul& cm, bool jc, float qiqj, float dx,
                                             float dy, float dz, float invR2, MixedAcc& a) {
  float E, W;
  if (LJ) {   // see pair_body
    float sr2 = float(p.c2) * invR2, sr6 = sr2 * sr2 * sr2;
    W = float(p.c1) * sr6 * (sr6 + sr6 - 1.0f);
    if (ENERGY) a.Ep += float(p.c0) * sr6 * (sr6 - 1.0f) + float(p.eshift);
    a.Wp += W;
    float g = W * invR2;
    a.fx += g * dx;
    a.fy += g * dy;
    a.fz += g * dz;
    return;
  }
  pair_model<float>(p, invR2, E, W);
  if (p.modifier != MOD_NONE) {
    float invR = (p.modifier == MOD_SHIFTED_FORCE || p.modifier == MOD_SMOOTHED || p.modifier == MOD_SHIFTED_SMOOTHED)
                     ? sqrt_fast(invR2) : 0.0f;
    apply_modifier<float>(p.modifier, float(p.eshift), float(p.fshift), float(p.factor), float(p.Rm2fac), invR,
                          invR2, E, W);
  }
  if (ENERGY) a.Ep += E;
  a.Wp += W;
  float Wsum = W;
  if (jc) {
    float invR = sqrt_fast(invR2);
    coul_model<float, ENERGY>(cm, invR, invR2, E, W);
    apply_modifier<float>(cm.modifier, float(cm.eshift), float(cm.fshift), float(cm.factor), float(cm.Rm2fac), invR,
                          invR2, E, W);
    if (ENERGY) a.Ec += qiqj * E;
    float Wq = qiqj * W;
    a.Wc += Wq;
    Wsum = Wsum + Wq;
  }
  float g = Wsum * invR2;
  a.fx += g * dx;
  a.fy += g * dy;
  a.fz += g * dz;
}

// LJ: 1 for the one-type LJ path (lj_path); 2 for layers whose pair models
// are all lj_cut (no modifier or an energy shift) or none, with any Coulomb
// model: each list entry then reads a 16-byte record of FP32 coefficients
// (ljt: c0, c1, c2, eshift; zeros for pair none) and the pair's Coulomb
// factor (ljk: kCoul, or 0 without Coulomb) instead of the full model record,
// and evaluates the same terms as pair_body_mx (an uncharged pair adds zero
// Coulomb terms); 0 for every other layer.
template <bool ENERGY, bool UNIFORM, int LJ, int TPA>
__global__ void __launch_bounds__(BLOCK)
k_pairs_mx(double* __restrict__ F, double* __restrict__ partials, const int4* __restrict__ Q4,
           const double* __restrict__ R, const double4* __restrict__ Rs4, const int* __restrict__ nbr,
           const int* __restrict__ nnb, const float* __restrict__ qf, const DevPair* __restrict__ pairs,
           const DevCoul* __restrict__ coulp,
           int ntypes, MixedCut mc, int N, int Npad, double* __restrict__ eatom, const float4* __restrict__ ljt,
           const float* __restrict__ ljk) {
  int blk = xcd_block();
  int gt = blk * BLOCK + threadIdx.x;
  int i = TPA == 1 ? gt : gt / TPA, sub = TPA == 1 ? 0 : gt % TPA;
  double fx = 0, fy = 0, fz = 0, Ep = 0, Ec = 0, Wp = 0, Wc = 0;
  if (i < N) {
    const DevCoul cm = *coulp;
    const int4 qi = Q4[i];
    const float qqi = qf[i];
    const DevPair* prow = pairs + (qi.w & Q4_TYPE) * ntypes;
    const int trow = (qi.w & Q4_TYPE) * ntypes;
    DevPair pu;
    if (UNIFORM) pu = pairs[0];   // uniform address: scalar loads, no VGPRs
    const int n = nnb[i];
    auto one = [&](int j, const int4& qj, MixedAcc& a) {
      float dx = (float)(int)((unsigned int)qi.x - (unsigned int)qj.x);
      float dy = (float)(int)((unsigned int)qi.y - (unsigned int)qj.y);
      float dz = (float)(int)((unsigned int)qi.z - (unsigned int)qj.z);
      float r2 = dx * dx + dy * dy + dz * dz;
      bool in = r2 < mc.lo;
      if (!in && r2 < mc.hi) in = exact_inside(Rs4, i, j, mc.Rc2s);
      if (in) {
        float invR2 = mc.K * __builtin_amdgcn_rcpf(r2);
        if (LJ == 1) {
          pair_body_mx<ENERGY, true>(pu, cm, false, 0.0f, dx, dy, dz, invR2, a);
        } else if (LJ == 2) {
          const int t = trow + (qj.w & Q4_TYPE);
          const float4 c = ljt[t];
          const float qiqj = ljk[t] * qqi * qf[j];
          float sr2 = c.z * invR2, sr6 = sr2 * sr2 * sr2, sr12 = sr6 * sr6;
          float W = c.y * (sr12 + sr12 - sr6);
          if (ENERGY) a.Ep += c.x * (sr12 - sr6) + c.w;
          a.Wp += W;
          float invR = sqrt_fast(invR2), E, Wq;
          coul_model<float, ENERGY>(cm, invR, invR2, E, Wq);
          apply_modifier<float>(cm.modifier, float(cm.eshift), float(cm.fshift), float(cm.factor), float(cm.Rm2fac),
                                invR, invR2, E, Wq);
          if (ENERGY) a.Ec += qiqj * E;
          Wq = qiqj * Wq;
          a.Wc += Wq;
          float g = (W + Wq) * invR2;
          a.fx += g * dx;
          a.fy += g * dy;
          a.fz += g * dz;
        } else if (UNIFORM) {
          bool jc = qqi != 0.0f && pu.coulomb && qf[j] != 0.0f;
          float qiqj = jc ? float(pu.kCoul) * qqi * qf[j] : 0.0f;
          pair_body_mx<ENERGY, false>(pu, cm, jc, qiqj, dx, dy, dz, invR2, a);
        } else {
          const DevPair p = prow[qj.w & Q4_TYPE];
          bool jc = qqi != 0.0f && p.coulomb && qf[j] != 0.0f;
          float qiqj = jc ? float(p.kCoul) * qqi * qf[j] : 0.0f;
          pair_body_mx<ENERGY, false>(p, cm, jc, qiqj, dx, dy, dz, invR2, a);
        }
      }
    };
    auto flush = [&](const MixedAcc& a) {
      fx += (double)a.fx; fy += (double)a.fy; fz += (double)a.fz;
      Wp += (double)a.Wp; Wc += (double)a.Wc;
      if (ENERGY) { Ep += (double)a.Ep; Ec += (double)a.Ec; }
    };
    if (LJ == 1) {
      // Branch-free groups of four: every entry is evaluated and masked, and
      // only a group holding an entry in the cutoff band takes the FP64 test.
      const float c0 = float(pu.c0), c1 = float(pu.c1), c2 = float(pu.c2), es = float(pu.eshift);
      // lists are padded to a multiple of 4 with the atom itself, masked by
      // index (a real neighbour at zero separation is evaluated, and gives a
      // non-finite force as in the FP64 path)
      const int n4 = (n + 3) & ~3;
      auto group4 = [&](const int* j, const int4* q) {
        MixedAcc a = { 0, 0, 0, 0, 0, 0, 0 };
#pragma unroll
        for (int u = 0; u < 4; u++) {
          float dx = (float)(int)((unsigned int)qi.x - (unsigned int)q[u].x);
          float dy = (float)(int)((unsigned int)qi.y - (unsigned int)q[u].y);
          float dz = (float)(int)((unsigned int)qi.z - (unsigned int)q[u].z);
          float r2 = dx * dx + dy * dy + dz * dz;
          if (j[u] == i) r2 = mc.hi;
          bool in = r2 < mc.lo;
          if (r2 >= mc.lo && r2 < mc.hi) in = exact_inside_R(R, Npad, mc.L, i, j[u], mc.Rc2s);
          float invR2 = mc.K * __builtin_amdgcn_rcpf(r2);
          float sr2 = c2 * invR2, sr6 = sr2 * sr2 * sr2;
          float W = c1 * sr6 * (sr6 + sr6 - 1.0f);
          if (ENERGY) a.Ep += in ? c0 * sr6 * (sr6 - 1.0f) + es : 0.0f;
          a.Wp += in ? W : 0.0f;
          float g = in ? W * invR2 : 0.0f;
          a.fx += g * dx;
          a.fy += g * dy;
          a.fz += g * dz;
        }
        flush(a);
      };
      // three groups of four in flight per iteration; each group is still
      // summed in FP32 and flushed to FP64 on its own.  The indices of the
      // next twelve entries are loaded before this block's terms, so their
      // latency hides behind the arithmetic.
      int k = 0;
      int4v jn[3];
      if (n4 >= 12) {
#pragma unroll
        for (int u = 0; u < 3; u++) jn[u] = nb_packet(nbr, u, i, Npad);
      }
      for (; k + 12 <= n4; k += 12) {
        int j[12];
        int4 q[12];
#pragma unroll
        for (int u = 0; u < 3; u++) {
          j[4 * u] = jn[u].x; j[4 * u + 1] = jn[u].y; j[4 * u + 2] = jn[u].z; j[4 * u + 3] = jn[u].w;
        }
#pragma unroll
        for (int u = 0; u < 12; u++) q[u] = Q4[j[u]];
        if (k + 24 <= n4) {
#pragma unroll
          for (int u = 0; u < 3; u++) jn[u] = nb_packet(nbr, (k >> 2) + 3 + u, i, Npad);
        }
        group4(j, q);
        group4(j + 4, q + 4);
        group4(j + 8, q + 8);
      }
      for (; k < n4; k += 4) {
        const int4v jv = nb_packet(nbr, k >> 2, i, Npad);
        int j[4] = { jv.x, jv.y, jv.z, jv.w };
        int4 q[4];
#pragma unroll
        for (int u = 0; u < 4; u++) q[u] = Q4[j[u]];
        group4(j, q);
      }
    } else if (TPA == 1) {
      int k = 0;
      for (; k + 4 <= n; k += 4) {
        const int4v jv = nb_packet(nbr, k >> 2, i, Npad);
        const int j0 = jv.x, j1 = jv.y, j2 = jv.z, j3 = jv.w;
        int4 q0 = Q4[j0], q1 = Q4[j1], q2 = Q4[j2], q3 = Q4[j3];
        MixedAcc a = { 0, 0, 0, 0, 0, 0, 0 };
        one(j0, q0, a); one(j1, q1, a); one(j2, q2, a); one(j3, q3, a);
        flush(a);
      }
      MixedAcc a = { 0, 0, 0, 0, 0, 0, 0 };
      for (; k < n; k++) {
        int j = nbr[nb_at(k, i, Npad)];
        one(j, Q4[j], a);
      }
      flush(a);
    } else {
      int k = sub;
      for (; k + 3 * TPA < n; k += 4 * TPA) {
        int j0 = nbr[nb_at(k, i, Npad)], j1 = nbr[nb_at(k + TPA, i, Npad)];
        int j2 = nbr[nb_at(k + 2 * TPA, i, Npad)], j3 = nbr[nb_at(k + 3 * TPA, i, Npad)];
        int4 q0 = Q4[j0], q1 = Q4[j1], q2 = Q4[j2], q3 = Q4[j3];
        MixedAcc a = { 0, 0, 0, 0, 0, 0, 0 };
        one(j0, q0, a); one(j1, q1, a); one(j2, q2, a); one(j3, q3, a);
        flush(a);
      }
      MixedAcc a = { 0, 0, 0, 0, 0, 0, 0 };
      for (; k < n; k += TPA) {
        int j = nbr[nb_at(k, i, Npad)];
        one(j, Q4[j], a);
      }
      flush(a);
    }
  }
  if (TPA > 1) {
    fx = group_sum<TPA>(fx);
    fy = group_sum<TPA>(fy);
    fz = group_sum<TPA>(fz);
  }
  if (i < N && sub == 0) {
    F[i] = mc.unit * fx;
    F[Npad + i] = mc.unit * fy;
    F[2 * Npad + i] = mc.unit * fz;
  }
  if (ENERGY && eatom) {   // atom energies: half of every pair term of the atom
    double e = Ep + Ec;
    if (TPA > 1) e = group_sum<TPA>(e);
    if (i < N && sub == 0) eatom[i] = 0.5 * e;
  }
  double v[4] = { Ep, Ec, Wp, Wc };
  block_sum<4>(v);
  if (threadIdx.x == 0) {
    partials[4 * blk + 0] = 0.5 * v[0];
    partials[4 * blk + 1] = 0.5 * v[1];
    partials[4 * blk + 2] = 0.5 * v[2];
    partials[4 * blk + 3] = 0.5 * v[3];
  }
}

// ---- integrator -------------------------------------------------------------
// boost, EmDeeData.f90:864-896 (free atoms): P = CP*P + CF*F;
// kinetic_energies, EmDeeData.f90:900-922: 2KE_x = sum invMass*P_x^2.
template <bool KE>
__global__ void k_boost(double* __restrict__ P, double* __restrict__ partials, const double* __restrict__ F,
                        const double* __restrict__ invMass, const int* __restrict__ freeFlag,
                        double CP, double CF, int translate, int N, int Npad) {
  int blk = blockIdx.x;
  int s = blk * BLOCK + threadIdx.x;
  double v[3] = { 0, 0, 0 };
  if (s < N && freeFlag[s]) {
    double im = invMass[s];
#pragma unroll
    for (int x = 0; x < 3; x++) {
      double p = P[x * Npad + s];
      if (translate) { p = CP * p + CF * F[x * Npad + s]; P[x * Npad + s] = p; }
      if (KE) v[x] = im * (p * p);
    }
  }
  if (KE) {
    block_sum<3>(v);
    if (threadIdx.x == 0) { partials[3 * blk] = v[0]; partials[3 * blk + 1] = v[1]; partials[3 * blk + 2] = v[2]; }
  }
}
// The same kick for free atoms of one mass, two neighbouring slots per
// thread (16-byte loads and stores; the second slot is left alone past N).
template <bool KE>
__global__ void k_boost_u2(double* __restrict__ P, double* __restrict__ partials, const double* __restrict__ F,
                           double CP, double CF, int translate, int N, int Npad, double im) {
  int blk = blockIdx.x;
  int s = 2 * (blk * BLOCK + threadIdx.x);
  double v[3] = { 0, 0, 0 };
  if (s < N) {
    bool two = s + 1 < N;
#pragma unroll
    for (int x = 0; x < 3; x++) {
      double2* px = reinterpret_cast<double2*>(P + (size_t)x * Npad + s);
      double2 p = *px;
      if (translate) {
        double2 f = *reinterpret_cast<const double2*>(F + (size_t)x * Npad + s);
        p.x = CP * p.x + CF * f.x;
        if (two) p.y = CP * p.y + CF * f.y;
        *px = p;
      }
      if (KE) v[x] = im * (p.x * p.x) + (two ? im * (p.y * p.y) : 0.0);
    }
  }
  if (KE) {
    block_sum<3>(v);
    if (threadIdx.x == 0) { partials[3 * blk] = v[0]; partials[3 * blk + 1] = v[1]; partials[3 * blk + 2] = v[2]; }
  }
}
// move, EmDeeData.f90:823-860 (free atoms): R = CR*R + CP*P*invMass.
__global__ void k_displace(double* __restrict__ R, const double* __restrict__ P, const double* __restrict__ invMass,
                           const int* __restrict__ freeFlag, double CR, double CP, int N, int Npad) {
  int s = blockIdx.x * BLOCK + threadIdx.x;
  if (s >= N || !freeFlag[s]) return;
  double im = invMass[s];
#pragma unroll
  for (int x = 0; x < 3; x++) R[x * Npad + s] = CR * R[x * Npad + s] + CP * P[x * Npad + s] * im;
}
// The same move when every atom is free and of one mass, two slots per thread.
__global__ void k_displace_u2(double* __restrict__ R, const double* __restrict__ P, double CR, double CP, int N,
                              int Npad, double im) {
  int s = 2 * (blockIdx.x * BLOCK + threadIdx.x);
  if (s >= N) return;
  bool two = s + 1 < N;
#pragma unroll
  for (int x = 0; x < 3; x++) {
    double2* rx = reinterpret_cast<double2*>(R + (size_t)x * Npad + s);
    double2 r = *rx, p = *reinterpret_cast<const double2*>(P + (size_t)x * Npad + s);
    r.x = CR * r.x + CP * p.x * im;
    if (two) r.y = CR * r.y + CP * p.y * im;
    *rx = r;
  }
}
// k_displace_u2 that also forms pass A of the rebuild trigger for the new
// positions (k_disp_a: per-block maximum of |R - R0|^2 with the smallest
// original index holding it) and their scaled copies, so that the force
// evaluation that follows does not read R and R0 again.  Two slots per
// thread: part holds one (max, index) pair per block of 2 BLOCK slots.
__global__ void k_displace_u2_trig(double* __restrict__ R, const double* __restrict__ P, double CR, double CP,
                                   int N, int Npad, double im, const double* __restrict__ R0,
                                   const int* __restrict__ perm, double* __restrict__ part, Scaled sc) {
  __shared__ double smax[BLOCK / 64];
  __shared__ int sidx[BLOCK / 64];
  int s = 2 * (blockIdx.x * BLOCK + threadIdx.x);
  double v = -1.0;
  int idx = INT_MAX;
  if (s < N) {
    bool two = s + 1 < N;
    double2 rn[3], r0[3];
#pragma unroll
    for (int x = 0; x < 3; x++) {
      double2* rx = reinterpret_cast<double2*>(R + (size_t)x * Npad + s);
      double2 r = *rx, p = *reinterpret_cast<const double2*>(P + (size_t)x * Npad + s);
      r.x = CR * r.x + CP * p.x * im;
      if (two) r.y = CR * r.y + CP * p.y * im;
      *rx = r;
      rn[x] = r;
      r0[x] = *reinterpret_cast<const double2*>(R0 + (size_t)x * Npad + s);
    }
    {
#pragma clang fp contract(off)
      double dx = rn[0].x - r0[0].x, dy = rn[1].x - r0[1].x, dz = rn[2].x - r0[2].x;
      v = dx * dx + dy * dy + dz * dz;
      idx = perm[s];
      if (two) {
        double ex = rn[0].y - r0[0].y, ey = rn[1].y - r0[1].y, ez = rn[2].y - r0[2].y;
        double w = ex * ex + ey * ey + ez * ez;
        int j = perm[s + 1];
        if (w > v || (w == v && j < idx)) { v = w; idx = j; }
      }
    }
    scale_vals(sc, rn[0].x, rn[1].x, rn[2].x, s);
    if (two) scale_vals(sc, rn[0].y, rn[1].y, rn[2].y, s + 1);
  }
  for (int off = 32; off > 0; off >>= 1) {
    double ov = __shfl_down(v, off, 64);
    int oi = __shfl_down(idx, off, 64);
    if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
  }
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) { smax[wid] = v; sidx[wid] = idx; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < BLOCK / 64; w++)
      if (smax[w] > v || (smax[w] == v && sidx[w] < idx)) { v = smax[w]; idx = sidx[w]; }
    part[2 * blockIdx.x] = v;
    part[2 * blockIdx.x + 1] = (double)idx;
  }
}

// ---- shadow terms of EmDee_verlet_step, free atoms (EmDeeCode.f90:1167-1204)
// s0 = 2.5 R + dt/2 invMass (P - dt/2 F), stored by ORIGINAL atom index so
// that a list rebuild (which permutes the slots) inside the step keeps it.
__global__ void k_free_shadow_pre(double* __restrict__ s0, const double* __restrict__ R, const double* __restrict__ P,
                                  const double* __restrict__ F, const double* __restrict__ invMass,
                                  const int* __restrict__ freeFlag, const int* __restrict__ perm, double dt, int N,
                                  int Npad) {
  int s = blockIdx.x * BLOCK + threadIdx.x;
  if (s >= N || !freeFlag[s]) return;
  double dt_2 = 0.5 * dt, im = invMass[s];
  int a = perm[s];
  for (int x = 0; x < 3; x++)
    s0[x * (size_t)N + a] = 2.5 * R[x * Npad + s] + dt_2 * im * (P[x * Npad + s] - dt_2 * F[x * Npad + s]);
}

// partials of {Ks_t, 0, Us}
__global__ void k_free_shadow_post(double* __restrict__ part, const double* __restrict__ s0,
                                   const double* __restrict__ R, const double* __restrict__ P,
                                   const double* __restrict__ F, const double* __restrict__ invMass,
                                   const int* __restrict__ freeFlag, const int* __restrict__ perm, double dt, int N,
                                   int Npad) {
  int s = blockIdx.x * BLOCK + threadIdx.x;
  double v[3] = { 0.0, 0.0, 0.0 };
  if (s < N && freeFlag[s]) {
    double dt_2 = 0.5 * dt, im = invMass[s];
    int a = perm[s];
    double rdot[3], p[3], f[3];
    for (int x = 0; x < 3; x++) {
      p[x] = P[x * Npad + s];
      f[x] = F[x * Npad + s];
      rdot[x] = 2.5 * R[x * Npad + s] + dt * im * (p[x] + dt_2 * f[x]) - s0[x * (size_t)N + a];
    }
    v[0] = rdot[0] * p[0] + rdot[1] * p[1] + rdot[2] * p[2];
    v[2] = im * (f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
  }
  block_sum<3>(v);
  if (threadIdx.x == 0) for (int k = 0; k < 3; k++) part[3 * blockIdx.x + k] = v[k];
}

// ---- EmDee_rdf (EmDeeCode.f90:1281-1395): pair counts per bin over the
// current neighbor lists, each pair once (original index a < b), with the
// reference's distance arithmetic (Rs = invL*R, d - anint(d), no FMA so that
// the bin of every pair is the reference's).  Integer atomics: exact counts.
template <typename PairOn>
__device__ __forceinline__ void rdf_atom(unsigned long long* __restrict__ count, const double* __restrict__ R,
                                         const int* __restrict__ nbr, const int* __restrict__ nnb,
                                         const int* __restrict__ perm, const int* __restrict__ type, PairOn on,
                                         double invL, double Rc2, double binsByRc, int bins, int s, int Npad) {
  int a = perm[s], it = type[s] + 1;
  double ri[3] = { __dmul_rn(invL, R[s]), __dmul_rn(invL, R[Npad + s]), __dmul_rn(invL, R[2 * Npad + s]) };
  int n = nnb[s];
  for (int k = 0; k < n; k++) {
    int t = nbr[nb_at(k, s, Npad)];
    if (perm[t] <= a) continue;
    int jt = type[t] + 1;
    if (!on(it, jt)) continue;
    double r2 = 0.0;
    for (int x = 0; x < 3; x++) {
      double d = __dsub_rn(ri[x], __dmul_rn(invL, R[x * Npad + t]));
      d = __dsub_rn(d, round(d));
      r2 = __dadd_rn(r2, __dmul_rn(d, d));
    }
    if (r2 < Rc2) {
      int bin = (int)__dmul_rn(sqrt(r2), binsByRc);
      if (bin >= bins) continue;   // r rounds up to Rc (the reference would index past its last bin)
      int lo = min(it, jt) - 1, hi = max(it, jt) - 1;
      int sp = lo + (hi + 1) * hi / 2;   // symm1D - 1
      atomicAdd(&count[bin + (size_t)bins * sp], 1ULL);
    }
  }
}

__global__ void k_rdf(unsigned long long* __restrict__ count, const double* __restrict__ R,
                      const int* __restrict__ nbr, const int* __restrict__ nnb, const int* __restrict__ perm,
                      const int* __restrict__ type, const uint8_t* __restrict__ pairOn, int nt, double invL,
                      double Rc2, double binsByRc, int bins, int N, int Npad) {
  int s = blockIdx.x * BLOCK + threadIdx.x;
  if (s >= N) return;
  auto on = [pairOn, nt](int it, int jt) { return pairOn[it * (nt + 1) + jt] != 0; };
  rdf_atom(count, R, nbr, nnb, perm, type, on, invL, Rc2, binsByRc, bins, s, Npad);
}

// ---- EmDeeGPU_rdf_noalloc: the same counts and normalisation as EmDee_rdf,
// with every buffer in `scratch`, a file-scope device array (no allocation
// per call).  The requested type pairs travel as a bit mask and the
// per-pair normalisation as a table, both by value in the kernel arguments.
constexpr int RDF_NOALLOC_COUNTS = RDF_NOALLOC_MAX_COUNTS;   // bins x type-pair slots
constexpr int RDF_NOALLOC_VALUES = RDF_NOALLOC_MAX_VALUES;   // bins x requested pairs (the output g)
constexpr int RDF_NOALLOC_PAIRS = RDF_NOALLOC_MAX_PAIRS;     // requested pairs per call
constexpr int RDF_NOALLOC_TYPES = RDF_NOALLOC_MAX_TYPE;      // largest type index
// 1 MiB on every GPU the library runs on, held until the process ends.
static __device__ __align__(256) unsigned long long scratch[RDF_NOALLOC_COUNTS + RDF_NOALLOC_VALUES];

struct RdfMask { unsigned int row[RDF_NOALLOC_TYPES + 1]; };   // bit jt of row[it]: pair (it, jt) requested
struct RdfNorm {
  int slot[RDF_NOALLOC_PAIRS];          // symm1D(i, j) - 1 of requested pair p
  int same[RDF_NOALLOC_PAIRS];          // i == j
  double NiNj[RDF_NOALLOC_PAIRS];       // (double) Ni * (double) Nj
};

__global__ void k_fill_u64(unsigned long long* __restrict__ p, int n) {
  int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) p[i] = 0ULL;
}

__global__ void k_rdf_masked(unsigned long long* __restrict__ count, const double* __restrict__ R,
                             const int* __restrict__ nbr, const int* __restrict__ nnb, const int* __restrict__ perm,
                             const int* __restrict__ type, RdfMask mask, double invL, double Rc2, double binsByRc,
                             int bins, int N, int Npad) {
  int s = blockIdx.x * BLOCK + threadIdx.x;
  if (s >= N) return;
  auto on = [&mask](int it, int jt) {
    return it <= RDF_NOALLOC_TYPES && jt <= RDF_NOALLOC_TYPES && ((mask.row[it] >> jt) & 1u) != 0;
  };
  rdf_atom(count, R, nbr, nnb, perm, type, on, invL, Rc2, binsByRc, bins, s, Npad);
}

// g(bin, p) in the arithmetic order of EmDee_rdf's host normalisation
// (EmDeeCode.f90:1336-1350): count / norm / (3 b (b - 1) + 1), doubled for
// like pairs, over Ni Nj.
__global__ void k_rdf_normalise(double* __restrict__ g, const unsigned long long* __restrict__ count, RdfNorm t,
                                double norm, int bins, int pairs) {
  int idx = blockIdx.x * BLOCK + threadIdx.x;
  if (idx >= bins * pairs) return;
  int b = idx % bins, p = idx / bins;
  double r = __ddiv_rn(__ddiv_rn((double)(long long)count[b + (size_t)bins * t.slot[p]], norm),
                       (double)(3 * (b + 1) * b + 1));
  g[idx] = __ddiv_rn(t.same[p] ? __dmul_rn(2.0, r) : r, t.NiNj[p]);
}

// ------------------------------------------------------------ host side

void launch_check(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) fatal("HIP kernel launch", std::string(what) + ": " + hipGetErrorString(e));
}

// Device address of `scratch` on the current device.  Each GPU holds its own
// copy of the array at its own address, looked up once per device (with that
// device current) by hipGetSymbolAddress; the host-side name `scratch` is not
// a device address and is never passed to a kernel or a copy.
namespace {
constexpr int MAX_DEVICES = 64;
std::mutex scratchMutex;
void* scratchAddress[MAX_DEVICES] = {};
}  // namespace

static void* scratch_address() {
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) fatal("system setup", "device index out of range");
  std::lock_guard<std::mutex> lock(scratchMutex);
  if (!scratchAddress[dev]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      fatal("system setup", std::string("no device address for the scratch array: ") + hipGetErrorString(e));
    scratchAddress[dev] = ptr;
  }
  return scratchAddress[dev];
}

Device* dev_create(const DeviceSetup& s) {
  Device* d = new Device();
  d->scratch = scratch_address();
  d->N = s.N;
  d->Npad = div_up(std::max(s.N, 1), BLOCK) * BLOCK;
  d->ntypes = s.ntypes;
  d->nlayers = s.nlayers;
  int Np = d->Npad;
  HIP_CHECK(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
  HIP_CHECK(hipEventCreate(&d->ev0));
  HIP_CHECK(hipEventCreate(&d->ev1));
  d->R.alloc(3 * (size_t)Np); d->P.alloc(3 * (size_t)Np); d->Rs4.alloc(Np); d->R0.alloc(3 * (size_t)Np);
  d->F.alloc(3 * (size_t)Np * s.nlayers);
  d->invMass.alloc(Np); d->charge.alloc(Np);
  d->type.alloc(Np); d->body.alloc(Np); d->perm.alloc(Np); d->freeFlag.alloc(Np); d->chargedFlag.alloc(Np);
  d->tmp3.alloc(3 * (size_t)Np); d->tmpF.alloc(3 * (size_t)Np * s.nlayers); d->tmp1.alloc(Np); d->tmpI.alloc(Np);
  d->tmpP.alloc(3 * (size_t)Np); d->tmpQ.alloc(Np);
  d->tmpBody.alloc(Np); d->tmpPerm.alloc(Np); d->tmpFree.alloc(Np); d->tmpCharged.alloc(Np);
  d->stage.alloc(3 * (size_t)std::max(s.N, 1));
  HIP_CHECK(hipHostMalloc(&d->hstage, 3 * sizeof(double) * std::max(s.N, 1), hipHostMallocDefault));
  HIP_CHECK(hipHostMalloc(&d->hresult, 64 * sizeof(double), hipHostMallocDefault));
  HIP_CHECK(hipHostMalloc((void**)&d->hpost, 64 * sizeof(double), hipHostMallocCoherent));
  HIP_CHECK(hipHostMalloc((void**)&d->hseq, 64, hipHostMallocCoherent));
  HIP_CHECK(hipHostGetDevicePointer((void**)&d->hpostDev, d->hpost, 0));
  HIP_CHECK(hipHostGetDevicePointer((void**)&d->hseqDev, d->hseq, 0));
  *d->hseq = 0;
  d->result.alloc(64);
  d->partials.alloc(64 * (size_t)(Np / BLOCK + 1));   // up to 16 threads per atom x 4 sums
  d->trigPart.alloc(2 * (size_t)(Np / BLOCK + 1));
  d->trig.alloc(4);
  HIP_CHECK(hipMemsetAsync(d->Rs4.p, 0, sizeof(double4) * Np, d->stream));
  d->key.alloc(Np); d->order.alloc(Np); d->nnb.alloc(Np); d->flags.alloc(4);
  HIP_CHECK(hipMemsetAsync(d->R.p, 0, 3 * sizeof(double) * Np, d->stream));
  HIP_CHECK(hipMemsetAsync(d->P.p, 0, 3 * sizeof(double) * Np, d->stream));
  HIP_CHECK(hipMemsetAsync(d->R0.p, 0, 3 * sizeof(double) * Np, d->stream));
  HIP_CHECK(hipMemsetAsync(d->F.p, 0, 3 * sizeof(double) * Np * s.nlayers, d->stream));
  HIP_CHECK(hipMemsetAsync(d->nnb.p, 0, sizeof(int) * Np, d->stream));
  HIP_CHECK(hipMemsetAsync(d->freeFlag.p, 0, sizeof(int) * Np, d->stream));
  int g = div_up(s.N, BLOCK);
  d->iperm.alloc(Np);
  if (s.N > 0) {
    hipLaunchKernelGGL(k_iota, dim3(g), dim3(BLOCK), 0, d->stream, d->perm.p, s.N);
    hipLaunchKernelGGL(k_iota, dim3(g), dim3(BLOCK), 0, d->stream, d->iperm.p, s.N);
  }
  launch_check("iota");
  // static per-atom data, initial order = original order
  std::vector<double> im(s.N), q(s.N);
  std::vector<int> free_(s.N), chg(s.N);
  for (int i = 0; i < s.N; i++) { im[i] = s.invMass[i]; q[i] = s.charge[i]; free_[i] = s.freeFlag[i]; chg[i] = std::fabs(s.charge[i]) > 2.220446049250313e-16; }
  d->uniformFree = s.N > 0;
  for (int i = 0; i < s.N; i++) d->uniformFree = d->uniformFree && free_[i] == 1 && im[i] == im[0];
  d->nfree = 0;
  for (int i = 0; i < s.N; i++) d->nfree += free_[i] != 0;
  d->invMass0 = s.N > 0 ? im[0] : 1.0;
  HIP_CHECK(hipMemcpyAsync(d->type.p, s.type.data(), sizeof(int) * s.N, hipMemcpyHostToDevice, d->stream));
  HIP_CHECK(hipMemcpyAsync(d->body.p, s.body.data(), sizeof(int) * s.N, hipMemcpyHostToDevice, d->stream));
  HIP_CHECK(hipMemcpyAsync(d->invMass.p, im.data(), sizeof(double) * s.N, hipMemcpyHostToDevice, d->stream));
  HIP_CHECK(hipMemcpyAsync(d->charge.p, q.data(), sizeof(double) * s.N, hipMemcpyHostToDevice, d->stream));
  HIP_CHECK(hipMemcpyAsync(d->freeFlag.p, free_.data(), sizeof(int) * s.N, hipMemcpyHostToDevice, d->stream));
  HIP_CHECK(hipMemcpyAsync(d->chargedFlag.p, chg.data(), sizeof(int) * s.N, hipMemcpyHostToDevice, d->stream));
  d->hostType = s.type;
  d->pairs.alloc((size_t)s.nlayers * s.ntypes * s.ntypes);
  d->couls.alloc(s.nlayers);
  d->ljt.alloc((size_t)s.nlayers * s.ntypes * s.ntypes);
  d->ljk.alloc((size_t)s.nlayers * s.ntypes * s.ntypes);
  d->ljt64.alloc((size_t)s.nlayers * s.ntypes * s.ntypes);
  d->ljk64.alloc((size_t)s.nlayers * s.ntypes * s.ntypes);
  d->ljcOk.assign(s.nlayers, 0);
  d->interact.alloc((size_t)s.ntypes * s.ntypes);
  HIP_CHECK(hipMemsetAsync(d->interact.p, 0, s.ntypes * s.ntypes, d->stream));
  d->exOff.alloc(s.N + 1);
  d->exItem.alloc(1);
  HIP_CHECK(hipStreamSynchronize(d->stream));
  return d;
}

void dev_destroy(Device* d) {
  if (!d) return;
  (void)hipStreamSynchronize(d->stream);
  for (DBuf<double>* b : { &d->R, &d->P, &d->R0, &d->invMass, &d->charge, &d->F, &d->tmp3, &d->tmpF, &d->tmp1, &d->tmpP, &d->tmpQ, &d->stage, &d->partials, &d->trigPart, &d->result, &d->trig, &d->shadow }) b->release();
  d->Rs4.release(); d->Q4.release(); d->P16.release(); d->qf.release(); d->eatom.release();
  for (DBuf<int>* b : { &d->type, &d->body, &d->perm, &d->freeFlag, &d->chargedFlag, &d->tmpI, &d->tmpBody, &d->tmpPerm, &d->tmpFree, &d->tmpCharged, &d->key, &d->cellCount, &d->cellStart, &d->order, &d->blockSums, &d->nbr, &d->nnb, &d->flags, &d->exOff, &d->exItem, &d->iperm }) b->release();
  d->pairs.release(); d->couls.release(); d->interact.release(); d->ljt.release(); d->ljk.release(); d->ljt64.release();
  d->ljk64.release();
  dev_ewald_release(d);
  dev_structs_release(d);
  if (d->hstage) (void)hipHostFree(d->hstage);
  if (d->hresult) (void)hipHostFree(d->hresult);
  if (d->hpost) (void)hipHostFree(d->hpost);
  if (d->hseq) (void)hipHostFree(d->hseq);
  (void)hipEventDestroy(d->ev0); (void)hipEventDestroy(d->ev1);
  (void)hipStreamDestroy(d->stream);
  delete d;
}

void dev_set_charges(Device* d, const std::vector<double>& q) {
  int N = d->N;
  std::vector<int> chg(N);
  for (int i = 0; i < N; i++) chg[i] = std::fabs(q[i]) > 2.220446049250313e-16;
  d->anyCharged = std::any_of(chg.begin(), chg.end(), [](int c) { return c != 0; });
  // original order -> sorted order through perm
  DBuf<double> tq; tq.alloc(N);
  DBuf<int> tc; tc.alloc(N);
  HIP_CHECK(hipMemcpyAsync(tq.p, q.data(), sizeof(double) * N, hipMemcpyHostToDevice, d->stream));
  HIP_CHECK(hipMemcpyAsync(tc.p, chg.data(), sizeof(int) * N, hipMemcpyHostToDevice, d->stream));
  int g = div_up(N, BLOCK);
  hipLaunchKernelGGL(k_gather_scalar<double>, dim3(g), dim3(BLOCK), 0, d->stream, d->charge.p, tq.p, d->perm.p, N);
  hipLaunchKernelGGL(k_gather_scalar<int>, dim3(g), dim3(BLOCK), 0, d->stream, d->chargedFlag.p, tc.p, d->perm.p, N);
  launch_check("set_charges");
  HIP_CHECK(hipStreamSynchronize(d->stream));
  tq.release(); tc.release();
}

void dev_set_exclusions(Device* d, const std::vector<int>& offset, const std::vector<int>& items) {
  d->nexcl = (int)items.size();
  d->exOff.alloc(offset.size());
  d->exItem.alloc(std::max<size_t>(items.size(), 1));
  HIP_CHECK(hipMemcpyAsync(d->exOff.p, offset.data(), sizeof(int) * offset.size(), hipMemcpyHostToDevice, d->stream));
  if (!items.empty())
    HIP_CHECK(hipMemcpyAsync(d->exItem.p, items.data(), sizeof(int) * items.size(), hipMemcpyHostToDevice, d->stream));
  HIP_CHECK(hipStreamSynchronize(d->stream));
}

void dev_set_interact(Device* d, const std::vector<uint8_t>& table) {
  d->allInteract = std::all_of(table.begin(), table.end(), [](uint8_t v) { return v != 0; });
  HIP_CHECK(hipMemcpyAsync(d->interact.p, table.data(), table.size(), hipMemcpyHostToDevice, d->stream));
  HIP_CHECK(hipStreamSynchronize(d->stream));
}

void dev_set_models(Device* d, int layer, const std::vector<DevPair>& pairs, const DevCoul& coul) {
  size_t nt2 = (size_t)d->ntypes * d->ntypes;
  d->hostPairs.resize(nt2 * d->nlayers);
  std::copy(pairs.begin(), pairs.begin() + nt2, d->hostPairs.begin() + nt2 * layer);
  HIP_CHECK(hipMemcpyAsync(d->pairs.p + nt2 * layer, pairs.data(), sizeof(DevPair) * nt2, hipMemcpyHostToDevice, d->stream));
  HIP_CHECK(hipMemcpyAsync(d->couls.p + layer, &coul, sizeof(DevCoul), hipMemcpyHostToDevice, d->stream));
  // compact FP32 records for the mixed kernels when every pair model of the
  // layer is lj_cut (no modifier or an energy shift) or none
  // (and their FP64 originals for the FP64 kernels)
  std::vector<float4> t(nt2);
  std::vector<float> k(nt2);
  std::vector<double4> t64(nt2);
  std::vector<double> k64(nt2);
  bool ok = true;
  for (size_t q = 0; q < nt2; q++) {
    const DevPair& p = pairs[q];
    bool lj = p.kind == PAIR_LJ_CUT, none = p.kind == PAIR_NONE;
    ok = ok && (lj || none) && (p.modifier == MOD_NONE || p.modifier == MOD_SHIFTED);
    double es = p.modifier == MOD_SHIFTED ? p.eshift : 0.0;
    t64[q] = lj ? make_double4(p.c0, p.c1, p.c2, es) : make_double4(0.0, 0.0, 0.0, es);
    k64[q] = p.coulomb ? p.kCoul : 0.0;
    t[q] = make_float4((float)t64[q].x, (float)t64[q].y, (float)t64[q].z, (float)t64[q].w);
    k[q] = (float)k64[q];
  }
  d->ljcOk[layer] = ok;
  HIP_CHECK(hipMemcpyAsync(d->ljt.p + nt2 * layer, t.data(), sizeof(float4) * nt2, hipMemcpyHostToDevice, d->stream));
  HIP_CHECK(hipMemcpyAsync(d->ljk.p + nt2 * layer, k.data(), sizeof(float) * nt2, hipMemcpyHostToDevice, d->stream));
  HIP_CHECK(hipMemcpyAsync(d->ljt64.p + nt2 * layer, t64.data(), sizeof(double4) * nt2, hipMemcpyHostToDevice,
                           d->stream));
  HIP_CHECK(hipMemcpyAsync(d->ljk64.p + nt2 * layer, k64.data(), sizeof(double) * nt2, hipMemcpyHostToDevice,
                           d->stream));
  HIP_CHECK(hipStreamSynchronize(d->stream));
}

static void upload_aos(Device* d, double* dst, const double* host, const int* mask) {
  int N = d->N;
  memcpy(d->hstage, host, 3 * sizeof(double) * N);
  HIP_CHECK(hipMemcpyAsync(d->stage.p, d->hstage, 3 * sizeof(double) * N, hipMemcpyHostToDevice, d->stream));
  hipLaunchKernelGGL(k_gather_aos, dim3(div_up(N, BLOCK)), dim3(BLOCK), 0, d->stream, dst, d->stage.p, d->perm.p, mask, N, d->Npad);
  launch_check("gather_aos");
  HIP_CHECK(hipStreamSynchronize(d->stream));
}
static void download_aos(Device* d, double* host, const double* src) {
  int N = d->N;
  hipLaunchKernelGGL(k_scatter_aos, dim3(div_up(N, BLOCK)), dim3(BLOCK), 0, d->stream, d->stage.p, src, d->perm.p, N, d->Npad);
  launch_check("scatter_aos");
  HIP_CHECK(hipMemcpyAsync(d->hstage, d->stage.p, 3 * sizeof(double) * N, hipMemcpyDeviceToHost, d->stream));
  HIP_CHECK(hipStreamSynchronize(d->stream));
  memcpy(host, d->hstage, 3 * sizeof(double) * N);
}

void dev_upload_positions(Device* d, const double* R) {
  d->posWrites++;
  if (d->N) upload_aos(d, d->R.p, R, nullptr);
}
void dev_download_positions(Device* d, double* R) { if (d->N) download_aos(d, R, d->R.p); }
void dev_upload_momenta(Device* d, const double* P) { if (d->N) upload_aos(d, d->P.p, P, d->freeFlag.p); }
void dev_download_momenta(Device* d, double* P) {
  int N = d->N;
  if (!N) return;
  hipLaunchKernelGGL(k_scatter_aos, dim3(div_up(N, BLOCK)), dim3(BLOCK), 0, d->stream, d->stage.p, d->P.p, d->perm.p, N, d->Npad);
  launch_check("scatter_aos");
  dev_bodies_particle_momenta(d);
  HIP_CHECK(hipMemcpyAsync(d->hstage, d->stage.p, 3 * sizeof(double) * N, hipMemcpyDeviceToHost, d->stream));
  HIP_CHECK(hipStreamSynchronize(d->stream));
  memcpy(P, d->hstage, 3 * sizeof(double) * N);
}
void dev_upload_forces(Device* d, int layer, const double* F) {
  if (d->N) upload_aos(d, d->F.p + 3 * (size_t)d->Npad * layer, F, nullptr);
}
void dev_download_forces(Device* d, int layer, double* F) {
  if (d->N) download_aos(d, F, d->F.p + 3 * (size_t)d->Npad * layer);
}
void dev_add_forces(Device* d, int layer, const double* F) {
  int N = d->N;
  if (!N) return;
  memcpy(d->hstage, F, 3 * sizeof(double) * N);
  HIP_CHECK(hipMemcpyAsync(d->stage.p, d->hstage, 3 * sizeof(double) * N, hipMemcpyHostToDevice, d->stream));
  hipLaunchKernelGGL(k_add_aos, dim3(div_up(N, BLOCK)), dim3(BLOCK), 0, d->stream, d->F.p + 3 * (size_t)d->Npad * layer, d->stage.p, d->perm.p, N, d->Npad);
  launch_check("add_aos");
  HIP_CHECK(hipStreamSynchronize(d->stream));
}

// The LJ-specialized pair kernels serve the layer: one type, no charges,
// pair_lj_cut with no modifier or an energy shift, and enough atoms to fill
// the chip with one thread per atom (smaller systems take the generic kernels,
// which split each list over 4 or 16 lanes).
static bool lj_path(const Device* d, int layer) {
  if (d->N < (1 << 17)) return false;
  if (d->ntypes != 1 || d->anyCharged || d->hostPairs.size() <= (size_t)layer) return false;
  const DevPair& hp = d->hostPairs[(size_t)layer];
  return hp.kind == PAIR_LJ_CUT && (hp.modifier == MOD_NONE || hp.modifier == MOD_SHIFTED);
}

void dev_trigger_launch(Device* d, double L, bool mixed, int layer) {
  int N = d->N;
  if (N == 0) return;
  int g = div_up(N, BLOCK);
  if (mixed) { d->Q4.alloc(d->Npad); d->qf.alloc(d->Npad); }
  // the mixed LJ kernel reads fixed-point positions only (its FP64 band
  // test divides R by L itself)
  bool needRs4 = !mixed || !lj_path(d, layer);
  // (the FP32 charges change only with the slot order: the list build writes them)
  int nb = g;
  if (d->trigPreAt == d->posWrites && d->trigPreL == L && d->trigPreMixed == mixed && d->trigPreRs4 == needRs4) {
    nb = d->trigPreBlocks;               // the drift formed pass A for these positions
  } else {
    Scaled sc{ needRs4 ? d->Rs4.p : nullptr, mixed ? d->Q4.p : nullptr, nullptr, d->type.p, d->charge.p,
               d->chargedFlag.p, L };
    hipLaunchKernelGGL(k_disp_a, dim3(g), dim3(BLOCK), 0, d->stream, d->trigPart.p, d->R.p, d->R0.p, d->perm.p, N,
                       d->Npad, sc);
  }
  d->trigKnown = true;
  d->trigMixed = mixed;
  d->trigRs4 = needRs4;
  d->trigL = L;
  d->trigSeq = ++d->seq;
  hipLaunchKernelGGL(k_disp_a_reduce, dim3(1), dim3(BLOCK), 0, d->stream, d->trig.p, d->trigPart.p, nb, d->hpostDev,
                     d->hseqDev, d->trigSeq);
  launch_check("trigger");
}

// With next <= max, approach lies in [max, 4 max] (the computed sums too, up
// to a few roundings), so pass C, which finds `next`, runs only when max
// alone does not decide: max <= skin^2 < 4 max.  The decision is the one of
// the full formula in every case.
bool dev_trigger_result(Device* d, double skinSq) {
  if (d->N == 0) return false;
  post_wait(d, d->trigSeq);
  double maximum = d->hpost[0], approach;
  if (maximum > skinSq) {
    approach = maximum;                    // a lower bound, enough to decide
  } else if (d->hpost[1] != 0.0 && 4.0 * maximum * (1.0 + 1e-12) < skinSq) {
    approach = 4.0 * maximum;              // an upper bound, enough to decide
  } else {
    double next = maximum;
    if (d->hpost[1] != 0.0) {              // pass C: max over atoms before the first maximum
      int N = d->N, g = div_up(N, BLOCK);
      hipLaunchKernelGGL(k_disp_c, dim3(g), dim3(BLOCK), 0, d->stream, d->partials.p, d->R.p, d->R0.p, d->perm.p,
                         d->trig.p, N, d->Npad);
      unsigned long long seq = ++d->seq;
      hipLaunchKernelGGL(k_disp_c_reduce, dim3(1), dim3(BLOCK), 0, d->stream, d->trig.p, d->partials.p, g,
                         d->hpostDev, d->hseqDev, seq);
      launch_check("trigger");
      post_wait(d, seq);
      next = d->hpost[2];
    }
    approach = maximum + 2 * std::sqrt(maximum * next) + next;
  }
  bool rebuild = approach > skinSq;
  // history for dev_trigger_unlikely; a rebuild starts the displacements anew
  d->trigPrev = rebuild ? 0.0 : d->trigLast;
  d->trigLast = rebuild ? 0.0 : approach;
  return rebuild;
}

// sqrt(approach) is the sum of two displacements and grows about linearly
// with time; a pair evaluation may be queued ahead of the trigger's answer
// when one and a half times the last step's growth still stays clear of the
// threshold.  A wrong guess costs one pair evaluation, never a result.
bool dev_trigger_unlikely(const Device* d, double skinSq) {
  if (d->trigLast < 0.0 || d->trigPrev < 0.0) return false;
  double s1 = std::sqrt(d->trigLast), s0 = std::sqrt(d->trigPrev);
  double p = s1 + 1.5 * std::max(s1 - s0, 0.0);
  return p * p < skinSq;
}


static void ensure_cells(Device* d, int M) {
  if (M == d->M) return;
  d->M = M;
  d->ncells = M * M * M;
  d->cellCount.alloc(d->ncells + 1);
  d->cellStart.alloc(d->ncells + 1);
  d->blockSums.alloc(div_up(d->ncells + 1, BLOCK) + 1);
}

void dev_rebuild(Device* d, double L, double xRc, double xInRc) {
  (void)xInRc;
  int N = d->N, Np = d->Npad;
  if (N == 0) return;
  d->posWrites++;                      // R and R0 in a new slot order
  int g = div_up(N, BLOCK);
  // lanes per atom of the list build: as many as keep at most 2^21 lanes in
  // all and at least four candidates per lane (the scanned region is ~2.6
  // times the list volume)
  double vol = 4.18879020478639 * xRc * xRc * xRc * ((double)N / (L * L * L));   // expected list length
  const bool simple = d->bodies.nb == 0 && d->nexcl == 0 && d->allInteract;
  const double cand = 2.6 * vol;
  const int S = (cand >= 16 * 4 && (long)N * 16 <= (1L << 21)) ? 16 : ((cand >= 4 * 4 && (long)N * 4 <= (1L << 21)) ? 4 : 1);
  int M = (int)std::floor(CELL_REACH * L / xRc);
  if (M < 2 * CELL_REACH + 1) M = 2 * CELL_REACH + 1;   // every cell of a row at most once
  if ((long)M * M * M > 64l * 1024 * 1024) M = 400;   // bound the cell table
  ensure_cells(d, M);
  int nc = d->ncells, nc1 = nc + 1, gc = div_up(nc1, BLOCK);
  // counting sort by cell: count, exclusive scan over ncells+1 entries
  // (start[ncells] = N), scatter, order inside cells
  HIP_CHECK(hipMemsetAsync(d->cellCount.p, 0, sizeof(int) * nc1, d->stream));
  hipLaunchKernelGGL(k_cell_key, dim3(g), dim3(BLOCK), 0, d->stream, d->key.p, d->cellCount.p, d->R.p, L, M, N, Np);
  hipLaunchKernelGGL(k_scan_block, dim3(gc), dim3(BLOCK), 0, d->stream, d->cellStart.p, d->blockSums.p, d->cellCount.p, nc1);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_T), 0, d->stream, d->blockSums.p, gc, (int*)nullptr);
  hipLaunchKernelGGL(k_scan_add, dim3(gc), dim3(BLOCK), 0, d->stream, d->cellStart.p, d->blockSums.p, nc1);
  HIP_CHECK(hipMemsetAsync(d->cellCount.p, 0, sizeof(int) * nc1, d->stream));
  // scattered (in-cell order of the atomics) into tmpI, which is free until
  // the permutation below, then ranked into order
  hipLaunchKernelGGL(k_cell_scatter, dim3(g), dim3(BLOCK), 0, d->stream, d->tmpI.p, d->cellCount.p, d->key.p, d->cellStart.p, N);
  hipLaunchKernelGGL(k_cell_rank, dim3(g), dim3(BLOCK), 0, d->stream, d->order.p, d->tmpI.p, d->key.p, d->cellStart.p,
                     d->perm.p, N);
  launch_check("cell sort");
  // every per-slot array into the new order (written to scratch, swapped),
  // R0 = R, and the inverse permutation
  {
    PermuteAll pa;
    pa.order = d->order.p;
    pa.R = d->R.p; pa.Rn = d->tmp3.p; pa.R0 = d->R0.p;
    pa.P = d->P.p; pa.Pn = d->tmpP.p;
    pa.F = d->nlayers > 1 ? d->F.p : nullptr; pa.Fn = d->tmpF.p; pa.nF = 3 * d->nlayers;
    pa.invMass = d->invMass.p; pa.invMassn = d->tmp1.p;
    pa.charge = d->charge.p; pa.chargen = d->tmpQ.p;
    pa.type = d->type.p; pa.typen = d->tmpI.p;
    pa.body = d->body.p; pa.bodyn = d->tmpBody.p;
    pa.perm = d->perm.p; pa.permn = d->tmpPerm.p;
    pa.freeFlag = d->freeFlag.p; pa.freeFlagn = d->tmpFree.p;
    pa.chargedFlag = d->chargedFlag.p; pa.chargedFlagn = d->tmpCharged.p;
    pa.iperm = d->iperm.p;
    hipLaunchKernelGGL(k_permute_all, dim3(g), dim3(BLOCK), 0, d->stream, pa, N, Np);
    auto swapD = [](DBuf<double>& x, DBuf<double>& y) { std::swap(x.p, y.p); std::swap(x.n, y.n); };
    auto swapI = [](DBuf<int>& x, DBuf<int>& y) { std::swap(x.p, y.p); std::swap(x.n, y.n); };
    swapD(d->R, d->tmp3); swapD(d->P, d->tmpP); swapD(d->invMass, d->tmp1); swapD(d->charge, d->tmpQ);
    if (d->nlayers > 1) swapD(d->F, d->tmpF);
    swapI(d->type, d->tmpI); swapI(d->body, d->tmpBody); swapI(d->perm, d->tmpPerm);
    swapI(d->freeFlag, d->tmpFree); swapI(d->chargedFlag, d->tmpCharged);
  }
  launch_check("permute");
  // scaled and fixed-point positions and keys in the new order (functions of
  // R, type and charge)
  d->Q4.alloc(Np);
  d->qf.alloc(Np);
  d->P16.alloc(Np + 16);               // the segment list build reads up to 16 slots past N
  hipLaunchKernelGGL(k_scale4q, dim3(g), dim3(BLOCK), 0, d->stream, d->Rs4.p, d->Q4.p, d->qf.p, d->R.p, d->type.p,
                     d->charge.p, d->chargedFlag.p, L, N, Np, d->P16.p, d->key.p, M);

  // Verlet list, growing the ELL capacity on overflow
  double xRc2 = (xRc * xRc) / (L * L);
  const double two32 = 4294967296.0, rc2u = xRc2 * two32 * two32, band = band_rel(xRc2);
  float lo = (float)(rc2u * (1.0 - band)), hi = (float)(rc2u * (1.0 + band));
  if (d->cap == 0) {
    d->cap = std::max(32, (int)(1.35 * vol + 16));
    d->cap = (d->cap + 7) & ~7;
  }
  for (;;) {
    d->nbr.alloc((size_t)d->cap * Np);
    HIP_CHECK(hipMemsetAsync(d->flags.p, 0, sizeof(int) * 4, d->stream));
    int gl = div_up((long)N * S, BLOCK);
    hipEvent_t e0 = d->profiling ? d->ev0 : nullptr, e1 = d->profiling ? d->ev1 : nullptr;
    // the segment-streaming form addresses list entries with 32-bit element
    // offsets (slot + k * Npad, k <= cap) and needs cells of edge <= 1/12 box
    if (simple && S == 1 && M >= 12 && (double)(d->cap + 1) * Np < 2147483648.0) {
      int lo16, hi16;
      band16(xRc2, &lo16, &hi16);
      hipExtLaunchKernelGGL(k_build_list_seg, dim3(gl), dim3(BLOCK), 0, d->stream, e0, e1, 0, d->nbr.p, d->nnb.p,
                            d->flags.p, d->partials.p, (const double4*)d->Rs4.p, (const int*)d->key.p,
                            (const int*)d->cellStart.p, M, xRc2, d->cap, N, Np, (const uint2*)d->P16.p, lo16, hi16);
    } else {
      auto kern = simple ? (S == 16 ? k_build_list<true, 16> : S == 4 ? k_build_list<true, 4> : k_build_list<true, 1>)
                         : (S == 16 ? k_build_list<false, 16> : S == 4 ? k_build_list<false, 4> : k_build_list<false, 1>);
      hipExtLaunchKernelGGL(kern, dim3(gl), dim3(BLOCK), 0, d->stream, e0, e1, 0, d->nbr.p, d->nnb.p, d->flags.p,
                            d->partials.p, (const double4*)d->Rs4.p, (const int*)d->key.p, (const int*)d->cellStart.p,
                            (const int*)d->body.p, (const int*)d->perm.p,
                            d->allInteract ? nullptr : (const uint8_t*)d->interact.p,
                            (const int*)d->exOff.p, (const int*)d->exItem.p, d->nexcl, d->ntypes, M, xRc2, d->cap, N, Np,
                            (const int4*)d->Q4.p, lo, hi);
    }
    launch_check("build_list");
    hipLaunchKernelGGL(k_reduce_partials<1>, dim3(1), dim3(BLOCK), 0, d->stream, d->result.p + 8, d->partials.p, gl);
    HIP_CHECK(hipMemcpyAsync(d->hresult + 8, d->result.p + 8, sizeof(double), hipMemcpyDeviceToHost, d->stream));
    int* hf = reinterpret_cast<int*>(d->hresult + 16);
    HIP_CHECK(hipMemcpyAsync(hf, d->flags.p, sizeof(int), hipMemcpyDeviceToHost, d->stream));
    HIP_CHECK(hipStreamSynchronize(d->stream));
    if (d->profiling) {
      float ms = 0; HIP_CHECK(hipEventElapsedTime(&ms, d->ev0, d->ev1));
      d->listMs += ms; d->listLaunches++;
    }
    if (hf[0] == 0) break;
    d->cap = ((int)(hf[0] * 1.2) + 15) & ~7;
  }
  d->listEntries = d->hresult[8];
  d->built = true;
  d->listVersion++;
}

int dev_neighbor_pairs(Device* d, int max, int* out) {
  int N = d->N, Np = d->Npad;
  std::vector<int> nnb(N), perm(N), nbr((size_t)d->cap * Np);
  HIP_CHECK(hipMemcpyAsync(nnb.data(), d->nnb.p, sizeof(int) * N, hipMemcpyDeviceToHost, d->stream));
  HIP_CHECK(hipMemcpyAsync(perm.data(), d->perm.p, sizeof(int) * N, hipMemcpyDeviceToHost, d->stream));
  HIP_CHECK(hipMemcpyAsync(nbr.data(), d->nbr.p, sizeof(int) * (size_t)d->cap * Np, hipMemcpyDeviceToHost, d->stream));
  HIP_CHECK(hipStreamSynchronize(d->stream));
  int n = 0;
  for (int s = 0; s < N; s++)
    for (int k = 0; k < nnb[s]; k++) {
      int t = nbr[nb_at(k, s, Np)];
      int a = perm[s] + 1, b = perm[t] + 1;
      if (a < b) {
        if (n < max) { out[2 * n] = a; out[2 * n + 1] = b; }
        n++;
      }
    }
  return n;
}

void dev_compute_pairs(Device* d, const PairCall& c, ForceSums* out) {
  int N = d->N, Np = d->Npad;
  double* F = d->F.p + 3 * (size_t)Np * c.layer;
  if (N == 0) { *out = ForceSums{ 0, 0, 0, 0 }; return; }
  int g = div_up(N, BLOCK);
  if (c.scaled) {
    // the trigger of this force evaluation (or the list build) wrote them
  } else if (c.mixed) {
    d->Q4.alloc(Np);
    d->qf.alloc(Np);
    hipLaunchKernelGGL(k_scale4q, dim3(g), dim3(BLOCK), 0, d->stream, d->Rs4.p, d->Q4.p, d->qf.p, d->R.p, d->type.p,
                       d->charge.p, d->chargedFlag.p, c.L, N, Np, (uint2*)nullptr, (int*)nullptr, 0);
  } else {
    hipLaunchKernelGGL(k_scale4, dim3(g), dim3(BLOCK), 0, d->stream, d->Rs4.p, d->R.p, d->type.p, c.L, N, Np);
  }
  const DevPair* pt = d->pairs.p + (size_t)c.layer * c.ntypes * c.ntypes;
  const DevCoul* cp = d->couls.p + c.layer;
  // with profiling on, the launch itself records the kernel's start and end
  // (hipExtLaunchKernelGGL), so the time excludes the host's launch latency
  hipEvent_t e0 = d->profiling ? d->ev0 : nullptr, e1 = d->profiling ? d->ev1 : nullptr;
  bool U = c.ntypes == 1;
  bool LJ = lj_path(d, c.layer);
  // threads per atom: enough wavefronts to fill the chip (at least 2^18
  // lanes) and about 24 list entries per lane, so that long lists (water at
  // rc 10 A holds ~700 entries per atom) are spread over many waves
  int tpa = 1;
  if (!LJ) {
    long lanes = std::max(1L << 18, (long)(d->listEntries / 24.0));
    long want = lanes / std::max(N, 1);
    tpa = want >= 16 ? 16 : (want >= 4 ? 4 : 1);
  }
  int gp = div_up((long)N * tpa, BLOCK);
  // the pair kernels write four sums per block into partials, which holds
  // 64 (Npad / BLOCK + 1) values: enough for up to 16 lanes per atom
  if ((size_t)gp * 4 > d->partials.n) fatal("compute_forces", "too many lanes per atom for the partial sums");
  double* ea = d->wantAtomE ? d->eatom.p : nullptr;
  int variant = (c.energy ? 4 : 0) + (U ? 2 : 0) + (LJ ? 1 : 0);
  if (c.mixed) {
    // cutoff in fixed-point units: Rc2 (scaled) * 2^64
    MixedCut mc;
    double two32 = 4294967296.0, rc2u = c.Rc2 * two32 * two32;
    mc.K = (float)(two32 * two32 / (c.L * c.L));
    mc.lo = (float)(rc2u * (1.0 - band_rel(c.Rc2)));
    mc.hi = (float)(rc2u * (1.0 + band_rel(c.Rc2)));
    mc.Rc2s = c.Rc2;
    mc.unit = c.L / two32;
    mc.L = c.L;
    const float4* ljt = d->ljt.p + (size_t)c.layer * c.ntypes * c.ntypes;
    const float* ljk = d->ljk.p + (size_t)c.layer * c.ntypes * c.ntypes;
    const bool ljc = d->ljcOk[c.layer] != 0;
#define LAUNCH_MX(E, U, J, T)                                                                                      \
  hipExtLaunchKernelGGL((k_pairs_mx<E, U, J, T>), dim3(gp), dim3(BLOCK), 0, d->stream, e0, e1, 0, F,                \
                        d->partials.p, (const int4*)d->Q4.p, (const double*)d->R.p, (const double4*)d->Rs4.p,     \
                        (const int*)d->nbr.p, (const int*)d->nnb.p, (const float*)d->qf.p, pt, cp, c.ntypes, mc, N, \
                        Np, ea, ljt, ljk)
#define GENERIC_MX_J(E, U, J)                                                                                      \
  do {                                                                                                            \
    if (tpa == 16) LAUNCH_MX(E, U, J, 16); else if (tpa == 4) LAUNCH_MX(E, U, J, 4);                              \
    else LAUNCH_MX(E, U, J, 1);                                                                                   \
  } while (0)
#define GENERIC_MX(E, U)                                                                                           \
  do {                                                                                                            \
    if (ljc) GENERIC_MX_J(E, U, 2); else GENERIC_MX_J(E, U, 0);                                                   \
  } while (0)
    switch (variant) {
      case 7: LAUNCH_MX(true, true, 1, 1); break;
      case 6: GENERIC_MX(true, true); break;
      case 4: GENERIC_MX(true, false); break;
      case 3: LAUNCH_MX(false, true, 1, 1); break;
      case 2: GENERIC_MX(false, true); break;
      default: GENERIC_MX(false, false); break;
    }
#undef GENERIC_MX
#undef GENERIC_MX_J
#undef LAUNCH_MX
  } else {
    const double4* ljt = d->ljt64.p + (size_t)c.layer * c.ntypes * c.ntypes;
    const double* ljk = d->ljk64.p + (size_t)c.layer * c.ntypes * c.ntypes;
    const bool ljc = d->ljcOk[c.layer] != 0;
#define LAUNCH_PAIRS(E, U, J, T)                                                                                    \
  hipExtLaunchKernelGGL((k_pairs<E, false, U, J, T>), dim3(gp), dim3(BLOCK), 0, d->stream, e0, e1, 0, F,           \
                        d->partials.p, (const double4*)d->Rs4.p, (const int*)d->nbr.p, (const int*)d->nnb.p,       \
                        (const double*)d->charge.p, (const int*)d->chargedFlag.p, pt, cp, c.ntypes, c.L, c.invL2,   \
                        c.Rc2, N, Np, ea, ljt, ljk)
#define GENERIC_J(E, U, J)                                                                                         \
  do {                                                                                                            \
    if (tpa == 16) LAUNCH_PAIRS(E, U, J, 16); else if (tpa == 4) LAUNCH_PAIRS(E, U, J, 4);                        \
    else LAUNCH_PAIRS(E, U, J, 1);                                                                                \
  } while (0)
#define GENERIC(E, U)                                                                                              \
  do {                                                                                                            \
    if (ljc) GENERIC_J(E, U, 2); else GENERIC_J(E, U, 0);                                                         \
  } while (0)
    switch (variant) {
      case 7: LAUNCH_PAIRS(true, true, 1, 1); break;
      case 6: GENERIC(true, true); break;
      case 4: GENERIC(true, false); break;
      case 3: LAUNCH_PAIRS(false, true, 1, 1); break;
      case 2: GENERIC(false, true); break;
      default: GENERIC(false, false); break;
    }
#undef GENERIC
#undef GENERIC_J
#undef LAUNCH_PAIRS
  }
  launch_check("pairs");
  // the pair sums have a slot of their own, so that a later post (the
  // kinetic energy of the kick that follows a deferred call) leaves them
  d->pairSeq = ++d->seq;
  hipLaunchKernelGGL(k_reduce_post<4>, dim3(1), dim3(BLOCK), 0, d->stream, d->hpostDev + POST_PAIRS, d->hseqDev,
                     d->pairSeq, d->partials.p, gp);
  launch_check("reduce");
  if (!c.defer) dev_pairs_result(d, out);
}

void dev_pairs_result(Device* d, ForceSums* out) {
  if (d->N == 0) { *out = ForceSums{ 0, 0, 0, 0 }; return; }
  post_wait(d, d->pairSeq);
  if (d->profiling) {
    float ms = 0;
    HIP_CHECK(hipEventSynchronize(d->ev1));
    HIP_CHECK(hipEventElapsedTime(&ms, d->ev0, d->ev1));
    d->pairMs += ms; d->pairLaunches++;
  }
  const double* h = d->hpost + POST_PAIRS;
  out->Epair = h[0];
  out->Ecoul = h[1];
  out->Wpair = h[2];
  out->Wcoul = h[3];
}

void dev_boost(Device* d, int layer, double CP, double CF, bool translate, double* twoKE) {
  int N = d->N, Np = d->Npad;
  // no free atoms (all in rigid bodies): nothing to kick, and their kinetic
  // energy is the sum of no terms
  if (N == 0 || d->nfree == 0) { if (twoKE) twoKE[0] = twoKE[1] = twoKE[2] = 0; return; }
  int g = div_up(N, BLOCK);
  const double* F = d->F.p + 3 * (size_t)Np * layer;
  const int g2 = div_up(div_up(N, 2), BLOCK);
  if (twoKE) {
    if (d->uniformFree) {
      hipLaunchKernelGGL(k_boost_u2<true>, dim3(g2), dim3(BLOCK), 0, d->stream, d->P.p, d->partials.p, F, CP, CF,
                         (int)translate, N, Np, d->invMass0);
      g = g2;
    } else {
      hipLaunchKernelGGL(k_boost<true>, dim3(g), dim3(BLOCK), 0, d->stream, d->P.p, d->partials.p, F, d->invMass.p,
                         d->freeFlag.p, CP, CF, (int)translate, N, Np);
    }
    unsigned long long seq = ++d->seq;
    hipLaunchKernelGGL(k_reduce_post<3>, dim3(1), dim3(BLOCK), 0, d->stream, d->hpostDev, d->hseqDev, seq, d->partials.p, g);
    launch_check("boost");
    post_wait(d, seq);
    for (int x = 0; x < 3; x++) twoKE[x] = d->hpost[x];
  } else if (translate) {
    if (d->uniformFree)
      hipLaunchKernelGGL(k_boost_u2<false>, dim3(g2), dim3(BLOCK), 0, d->stream, d->P.p, d->partials.p, F, CP, CF, 1, N,
                         Np, d->invMass0);
    else
      hipLaunchKernelGGL(k_boost<false>, dim3(g), dim3(BLOCK), 0, d->stream, d->P.p, d->partials.p, F, d->invMass.p,
                         d->freeFlag.p, CP, CF, 1, N, Np);
  }
  launch_check("boost");
}

void dev_kinetic(Device* d, double* twoKE) { dev_boost(d, 0, 1.0, 0.0, false, twoKE); }

void dev_displace(Device* d, double CR, double CP) {
  int N = d->N;
  if (N == 0 || d->nfree == 0) return;   // no free atom moves (rigid bodies move in dev_bodies_move)
  d->posWrites++;
  if (d->uniformFree && d->trigKnown) {
    // the drift with the next trigger's pass A, in the last trigger's set-up
    int g = div_up(div_up(N, 2), BLOCK);
    Scaled sc{ d->trigRs4 ? d->Rs4.p : nullptr, d->trigMixed ? d->Q4.p : nullptr, nullptr, d->type.p, d->charge.p,
               d->chargedFlag.p, d->trigL };
    hipLaunchKernelGGL(k_displace_u2_trig, dim3(g), dim3(BLOCK), 0, d->stream, d->R.p, d->P.p, CR, CP, N, d->Npad,
                       d->invMass0, d->R0.p, d->perm.p, d->trigPart.p, sc);
    d->trigPreAt = d->posWrites;
    d->trigPreBlocks = g;
    d->trigPreMixed = d->trigMixed;
    d->trigPreRs4 = d->trigRs4;
    d->trigPreL = d->trigL;
  } else if (d->uniformFree)
    hipLaunchKernelGGL(k_displace_u2, dim3(div_up(div_up(N, 2), BLOCK)), dim3(BLOCK), 0, d->stream, d->R.p, d->P.p, CR,
                       CP, N, d->Npad, d->invMass0);
  else
    hipLaunchKernelGGL(k_displace, dim3(div_up(N, BLOCK)), dim3(BLOCK), 0, d->stream, d->R.p, d->P.p, d->invMass.p,
                       d->freeFlag.p, CR, CP, N, d->Npad);
  launch_check("displace");
}

void dev_shadow_pre(Device* d, int layer, double dt, int mode) {
  int N = d->N, Np = d->Npad;
  int nb = d->bodies.nb;
  d->shadow.alloc(3 * (size_t)std::max(N, 1) + 7 * (size_t)nb);
  if (N > 0)
    hipLaunchKernelGGL(k_free_shadow_pre, dim3(div_up(N, BLOCK)), dim3(BLOCK), 0, d->stream, d->shadow.p, d->R.p, d->P.p,
                       d->F.p + 3 * (size_t)Np * layer, d->invMass.p, d->freeFlag.p, d->perm.p, dt, N, Np);
  launch_check("shadow pre");
  dev_bodies_shadow_pre(d, d->shadow.p + 3 * (size_t)std::max(N, 1), dt, mode);
}

void dev_shadow_post(Device* d, int layer, double dt, int mode, double* out3) {
  int N = d->N, Np = d->Npad;
  double bsum[3];
  dev_bodies_shadow_post(d, d->shadow.p + 3 * (size_t)std::max(N, 1), dt, mode, bsum);
  double fsum[3] = { 0.0, 0.0, 0.0 };
  if (N > 0) {
    int g = div_up(N, BLOCK);
    hipLaunchKernelGGL(k_free_shadow_post, dim3(g), dim3(BLOCK), 0, d->stream, d->partials.p, d->shadow.p, d->R.p,
                       d->P.p, d->F.p + 3 * (size_t)Np * layer, d->invMass.p, d->freeFlag.p, d->perm.p, dt, N, Np);
    launch_check("shadow post");
    hipLaunchKernelGGL(k_reduce_partials<3>, dim3(1), dim3(BLOCK), 0, d->stream, d->result.p + 40, d->partials.p, g);
    HIP_CHECK(hipMemcpyAsync(d->hresult + 40, d->result.p + 40, 3 * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    HIP_CHECK(hipStreamSynchronize(d->stream));
    for (int k = 0; k < 3; k++) fsum[k] = d->hresult[40 + k];
  }
  // body terms first, then free atoms, as the per-thread sums of the reference
  for (int k = 0; k < 3; k++) out3[k] = bsum[k] + fsum[k];
}

void dev_rdf(Device* d, int bins, double Rc, double L, const std::vector<uint8_t>& pairOn, int npt,
             std::vector<long long>& count) {
  int N = d->N;
  count.assign((size_t)bins * npt, 0);
  if (N == 0 || !d->built) return;
  DBuf<unsigned long long> c;
  DBuf<uint8_t> on;
  c.alloc((size_t)bins * npt);
  on.alloc(pairOn.size());
  HIP_CHECK(hipMemsetAsync(c.p, 0, sizeof(unsigned long long) * bins * npt, d->stream));
  HIP_CHECK(hipMemcpyAsync(on.p, pairOn.data(), pairOn.size(), hipMemcpyHostToDevice, d->stream));
  double invL = 1.0 / L, invL2 = invL * invL, Rc2 = Rc * Rc * invL2, binsByRc = bins / (Rc * invL);
  hipLaunchKernelGGL(k_rdf, dim3(div_up(N, BLOCK)), dim3(BLOCK), 0, d->stream, c.p, d->R.p, d->nbr.p, d->nnb.p,
                     d->perm.p, d->type.p, on.p, d->ntypes, invL, Rc2, binsByRc, bins, N, d->Npad);
  launch_check("rdf");
  std::vector<unsigned long long> h((size_t)bins * npt);
  HIP_CHECK(hipMemcpyAsync(h.data(), c.p, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost, d->stream));
  HIP_CHECK(hipStreamSynchronize(d->stream));
  for (size_t k = 0; k < h.size(); k++) count[k] = (long long)h[k];
  c.release();
  on.release();
}

// EmDee_rdf with no allocation: counts and results in `scratch` (through the
// address looked up at set-up), the pair mask and normalisation by value.
// The caller has checked the limits (RDF_NOALLOC_*).  Calls that overlap on
// one GPU share the array and race; each call clears the counts it uses.
void dev_rdf_noalloc(Device* d, int bins, double Rc, double L, const int* itype, const int* jtype, int pairs,
                     const double* NiNj, int npt, double norm, double* g) {
  int N = d->N;
  unsigned long long* count = static_cast<unsigned long long*>(d->scratch);
  double* gdev = reinterpret_cast<double*>(count + RDF_NOALLOC_COUNTS);
  RdfMask mask;
  RdfNorm t;
  for (int k = 0; k <= RDF_NOALLOC_TYPES; k++) mask.row[k] = 0u;
  for (int p = 0; p < RDF_NOALLOC_PAIRS; p++) { t.slot[p] = 0; t.same[p] = 0; t.NiNj[p] = 1.0; }
  for (int p = 0; p < pairs; p++) {
    int i = itype[p], j = jtype[p];
    mask.row[i] |= 1u << j;
    mask.row[j] |= 1u << i;
    int lo = std::min(i, j) - 1, hi = std::max(i, j) - 1;
    t.slot[p] = lo + (hi + 1) * hi / 2;
    t.same[p] = i == j;
    t.NiNj[p] = NiNj[p];
  }
  int nc = bins * npt;
  hipLaunchKernelGGL(k_fill_u64, dim3(div_up(nc, BLOCK)), dim3(BLOCK), 0, d->stream, count, nc);
  if (N > 0 && d->built) {
    double invL = 1.0 / L, invL2 = invL * invL, Rc2 = Rc * Rc * invL2, binsByRc = bins / (Rc * invL);
    hipLaunchKernelGGL(k_rdf_masked, dim3(div_up(N, BLOCK)), dim3(BLOCK), 0, d->stream, count, d->R.p, d->nbr.p,
                       d->nnb.p, d->perm.p, d->type.p, mask, invL, Rc2, binsByRc, bins, N, d->Npad);
  }
  hipLaunchKernelGGL(k_rdf_normalise, dim3(div_up(bins * pairs, BLOCK)), dim3(BLOCK), 0, d->stream, gdev, count, t,
                     norm, bins, pairs);
  launch_check("rdf (no allocation)");
  HIP_CHECK(hipMemcpyAsync(g, gdev, sizeof(double) * bins * pairs, hipMemcpyDeviceToHost, d->stream));
  HIP_CHECK(hipStreamSynchronize(d->stream));
}

void dev_set_atom_energies(Device* d, bool on) {
  d->wantAtomE = on;
  if (on) {
    d->eatom.alloc(d->Npad);
    HIP_CHECK(hipMemsetAsync(d->eatom.p, 0, sizeof(double) * d->Npad, d->stream));
  }
}
void dev_get_atom_energies(Device* d, double* host) {
  int N = d->N;
  std::vector<double> e(N);
  std::vector<int> perm(N);
  HIP_CHECK(hipMemcpyAsync(e.data(), d->eatom.p, sizeof(double) * N, hipMemcpyDeviceToHost, d->stream));
  HIP_CHECK(hipMemcpyAsync(perm.data(), d->perm.p, sizeof(int) * N, hipMemcpyDeviceToHost, d->stream));
  HIP_CHECK(hipStreamSynchronize(d->stream));
  for (int s = 0; s < N; s++) host[perm[s]] = e[s];
}

void dev_set_profiling(Device* d, bool on) { d->profiling = on; d->pairMs = d->listMs = 0; d->pairLaunches = d->listLaunches = 0; }
void dev_kernel_stats(Device* d, double* out, int n) {
  double v[7] = { d->pairMs, (double)d->pairLaunches, d->listMs, (double)d->listLaunches, (double)d->cap, (double)d->M,
                  d->listEntries };
  for (int k = 0; k < n && k < 7; k++) out[k] = v[k];
}
void dev_sync(Device* d) { HIP_CHECK(hipStreamSynchronize(d->stream)); }

// dst slot s holds original atom dperm[s], which sits at slot siperm[o] of src
__global__ void k_copy_phase(double* __restrict__ dR, double* __restrict__ dP, const double* __restrict__ sR,
                             const double* __restrict__ sP, const int* __restrict__ dperm,
                             const int* __restrict__ siperm, int N, int dNpad, int sNpad) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  int t = siperm[dperm[s]];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    dR[(size_t)c * dNpad + s] = sR[(size_t)c * sNpad + t];
    dP[(size_t)c * dNpad + s] = sP[(size_t)c * sNpad + t];
  }
}

void dev_copy_phase(Device* dst, Device* src) {
  if (dst->N != src->N) fatal("phase space sharing", "different numbers of atoms");
  dst->posWrites++;
  HIP_CHECK(hipStreamSynchronize(src->stream));
  if (dst->N > 0)
    hipLaunchKernelGGL(k_copy_phase, dim3(div_up(dst->N, BLOCK)), dim3(BLOCK), 0, dst->stream, dst->R.p, dst->P.p,
                       src->R.p, src->P.p, dst->perm.p, src->iperm.p, dst->N, dst->Npad, src->Npad);
  launch_check("phase copy");
  dev_bodies_copy(dst, src);
  HIP_CHECK(hipStreamSynchronize(dst->stream));
}

}  // namespace emdee
