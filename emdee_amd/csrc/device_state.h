// device_state.h -- HBM state of one system, shared by the kernel files
// (device.hip: neighbor lists, pair loop, integrator of free atoms;
// rigid.hip: rigid bodies).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "engine.h"

namespace emdee {

#define HIP_CHECK(x)                                                             \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess) fatal("HIP runtime", std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr int BLOCK = 256;
constexpr int NXCD = 8;

static inline int div_up(long a, long b) { return (int)((a + b - 1) / b); }

// Verlet list layout: each slot's entries in packets of four consecutive
// ints, one row of Npad packets per four entries, so entry k of slot s is at
// nb_at(k, s, Npad) and four entries of a slot are one 16-byte access (the
// capacity is a multiple of eight).
__host__ __device__ __forceinline__ size_t nb_at(int k, int s, int Npad) {
  return 4 * ((size_t)(k >> 2) * Npad + s) + (k & 3);
}
typedef int int4v __attribute__((ext_vector_type(4)));
// entries 4p .. 4p+3 of slot s
__device__ __forceinline__ int4v nb_packet(const int* __restrict__ nbr, int p, int s, int Npad) {
  return __builtin_nontemporal_load(reinterpret_cast<const int4v*>(nbr) + ((size_t)p * Npad + s));
}

// Blocks are dispatched round-robin over the 8 XCDs (block b runs on XCD
// b % 8).  Remap so that XCD x processes one contiguous range of slots: the
// neighbor positions a wave gathers then mostly live in that XCD's L2.
__device__ __forceinline__ int xcd_block() {
  int b = blockIdx.x, nb = gridDim.x;
  int q = nb / NXCD, rem = nb % NXCD, x = b % NXCD, r = b / NXCD;
  return x * q + (x < rem ? x : rem) + r;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Sums NV values over the block; thread 0 of the block gets the totals.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
  __shared__ double red[BLOCK / 64][NV];
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = wave_sum(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) red[wid][k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) {
      double s = 0;
      for (int w = 0; w < BLOCK / 64; w++) s += red[w][k];
      v[k] = s;
    }
  }
}

// Deterministic final reduction of per-block partials (fixed order).
template <int NV>
__global__ void k_reduce_partials(double* __restrict__ out, const double* __restrict__ partials, int nblocks) {
  double v[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = 0;
#pragma unroll 8
  for (int b = threadIdx.x; b < nblocks; b += BLOCK) {
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] += partials[NV * b + k];
  }
  block_sum<NV>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) out[k] = v[k];
  }
}

// The same reduction, with the NV totals written straight into pinned,
// coherent host memory followed by a sequence number: the host polls the
// number (post_wait) instead of copying the totals back and waiting on the
// stream, which saves a copy launch and the wake-up latency of a stream wait
// on every energy, kinetic-energy and list-trigger read.  (Finishing the
// reduction in the last block of the producing kernel instead would need a
// device-scope release fence per block, which writes back the XCD's L2 and
// costs more than the launch it saves.)
template <int NV>
__global__ void k_reduce_post(double* __restrict__ host, unsigned long long* __restrict__ hostSeq,
                              unsigned long long seq, const double* __restrict__ partials, int nblocks) {
  double v[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = 0;
#pragma unroll 8
  for (int b = threadIdx.x; b < nblocks; b += BLOCK) {
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] += partials[NV * b + k];
  }
  block_sum<NV>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) host[k] = v[k];
    __threadfence_system();
    *(volatile unsigned long long*)hostSeq = seq;
  }
}

// ------------------------------------------------------------ device state

// Slots in the posted-results block (64 doubles): the pair sums, the Ewald
// energy and virial, the learned model's energy, virial and overflow flag,
// the rigid bodies' kinetic-energy terms and their virial, the model's
// low-k energy and virial; the other posts use slot 0.
constexpr int POST_PAIRS = 8, POST_EWALD = 16, POST_MODEL = 24, POST_BODY_KE = 32, POST_BODY_W = 40, POST_LOWK = 48;

template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    if (count <= n && p) return;
    if (p) HIP_CHECK(hipFree(p));
    HIP_CHECK(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
    n = count;
  }
  void release() { if (p) (void)hipFree(p); p = nullptr; n = 0; }
};

// Rigid bodies (ArBee.f90 tBody), SoA over bodies; members in CSR order.
struct BodyState {
  int nb = 0, nm = 0;
  DBuf<int> start, orig;        // member range of body b: start[b]..start[b+1]-1; original atom index
  DBuf<double> mMass, d, delta; // member mass; body-frame and space-frame member offsets [3][nm]
  DBuf<double> mass, invMass, MoI, invMoI, coef;   // coef: I113, I313, I223, I212, m312 [5][nb]
  DBuf<double> rcm, pcm, q, pi, omega, F, tau;     // [3][nb] or [4][nb]
  DBuf<double> part;            // reduction partials
};

struct EwaldState;
struct StructState;

struct Device {
  int N = 0, Npad = 0, ntypes = 0, nlayers = 0;
  hipStream_t stream = nullptr;
  void* scratch = nullptr;   // device address of this GPU's `scratch` array (device.hip)
  // per-atom SoA arrays (sorted order)
  DBuf<double> R, P, R0, invMass, charge, F;   // F holds nlayers * 3 * Npad
  DBuf<double4> Rs4;      // {x/L, y/L, z/L, type}: one 32-byte gather per neighbor
  DBuf<int4> Q4;          // mixed form: 32-bit fixed-point scaled coordinates and type
  DBuf<uint2> P16;        // list build: 16-bit fixed-point scaled coordinates {x | y << 16, z | nonfinite << 31}
  DBuf<float> qf;         // mixed form: FP32 charge (0 when uncharged)
  DBuf<int> type, body, perm, freeFlag, chargedFlag;
  // permutation scratch, one per permuted array (swapped with the arrays)
  DBuf<double> tmp3, tmpF, tmp1, tmpP, tmpQ;
  DBuf<int> tmpI, tmpBody, tmpPerm, tmpFree, tmpCharged;
  // transfers
  DBuf<double> stage;     // 3N AoS, original order
  double* hstage = nullptr;  // pinned host mirror
  // cells
  int M = 0, ncells = 0;
  DBuf<int> key, cellCount, cellStart, order, blockSums;   // cellStart has ncells+1 entries
  // neighbor list
  int cap = 0;
  DBuf<int> nbr, nnb;
  DBuf<int> flags;        // [0] overflow max count
  bool built = false;
  long listVersion = 0;   // incremented at every list build
  // models
  DBuf<DevPair> pairs;    // nlayers * nt * nt
  std::vector<DevPair> hostPairs;   // host copy of the pair table
  bool anyCharged = false;
  DBuf<DevCoul> couls;    // nlayers
  DBuf<uint8_t> interact;
  DBuf<int> exOff, exItem;
  int nexcl = 0;
  bool allInteract = false;   // every type pair has a model (interaction table all ones)
  bool uniformFree = false;   // every atom free and of one mass: the integrator skips the per-atom reads
  int nfree = 0;              // atoms in no rigid body: none, and the kick and drift of free atoms are skipped
  double invMass0 = 1.0;      // that mass's inverse
  std::vector<int> hostType;
  // reductions
  DBuf<double> partials, result;
  double* hresult = nullptr;
  // posted results (k_reduce_post): coherent pinned host memory, its device
  // address, and the sequence number of the last post requested
  double* hpost = nullptr;
  double* hpostDev = nullptr;
  unsigned long long* hseq = nullptr;
  unsigned long long* hseqDev = nullptr;
  unsigned long long seq = 0;
  unsigned long long pairSeq = 0;   // post of the last pair evaluation (slot POST_PAIRS)
  // per layer: FP32 records {c0, c1, c2, eshift} and Coulomb factors of the
  // type pairs, used by the mixed kernels when ljcOk (k_pairs_mx, LJ = 2)
  DBuf<float4> ljt;
  DBuf<float> ljk;
  DBuf<double4> ljt64;   // the same records in FP64 for the FP64 kernels
  DBuf<double> ljk64;
  std::vector<char> ljcOk;
  unsigned long long trigSeq = 0;   // post of the last rebuild trigger (slot 0)
  // Every write of R on the device counts in posWrites.  The drift of
  // uniform free atoms forms the next trigger's pass A (partials, scaled
  // positions) on the way, with the set-up of the last trigger (trigMixed,
  // trigRs4, trigL); the trigger uses them while no other write intervened.
  unsigned long long posWrites = 0;
  bool trigKnown = false, trigMixed = false, trigRs4 = false;
  double trigL = 0.0;
  unsigned long long trigPreAt = ~0ull;   // posWrites when the drift formed pass A
  DBuf<double> trigPart;                  // pass A partials (no other reduction uses them)
  int trigPreBlocks = 0;
  bool trigPreMixed = false, trigPreRs4 = false;
  double trigPreL = 0.0;
  double trigLast = -1.0, trigPrev = -1.0;   // trigger values of the last two steps (-1: none)
  DBuf<double> trig;      // [0] max, [1] first original index holding it, [2] max before it
  // rigid bodies; iperm: original atom -> slot
  BodyState bodies;
  DBuf<int> iperm;
  // k-space (ewald.hip) and bonded terms (bonded.hip), created on demand
  DBuf<double> eatom;     // atom energies (slot order) when wantAtomE
  bool wantAtomE = false;
  DBuf<double> shadow;    // EmDee_verlet_step: s0 [3][N] (original order), then r0 [3][nb], q0 [4][nb]
  EwaldState* ewald = nullptr;
  StructState* structs = nullptr;
  // profiling
  bool profiling = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double pairMs = 0, listMs = 0;
  long pairLaunches = 0, listLaunches = 0;
  double listEntries = 0;
};


void launch_check(const char* what);
// Waits until the post with sequence number seq has arrived in pinned host
// memory (posts arrive in stream order).
void dev_post_wait(Device* d, unsigned long long seq);
void dev_ewald_release(Device* d);
void dev_structs_release(Device* d);
void dev_bodies_shadow_pre(Device* d, double* sh, double dt, int mode);
void dev_bodies_shadow_post(Device* d, const double* sh, double dt, int mode, double* out3);

}  // namespace emdee
