// surrogate.hip -- learned per-atom energy model that stands in for a costly
// part of the potential (the long-range electrostatics of a model layer).
// There is no counterpart in the reference.
//
// Model: for atom i of species s, the descriptor D_i holds, for every
// neighbour species t, NB radial Gaussians exp(-eta (r_ij - mu_k)^2) times the
// cosine cutoff 0.5 (cos(pi r/rc) + 1), summed over neighbours j of species t
// within rc (pairs that the engine's Verlet list holds: excluded and
// same-body pairs never appear).  The normalised descriptor feeds a per-species
// MLP  din -> H -> H -> 1  with SiLU activations; the atom energy is
// e_i = ystd_s * out + ymean_s and the model energy is the sum of e_i.
// Forces are -grad sum e_i: back-propagation through the MLP to dE/dD, then
// through the Gaussians to the pair separations.
//
// The two hidden-layer products and their transposes run on MFMA
// (v_mfma_f32_16x16x16_bf16: BF16 operands, FP32 accumulation); descriptors,
// activations and the pair back-propagation are FP32 with FP64 force
// accumulation.  Training fits per-atom energy labels (MSE) with Adam, with
// the weight gradients also on MFMA and deterministic partial reductions.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/emdee_gpu.h"
#include "device_state.h"
#include "kspace.h"
#include "surrogate.h"

namespace emdee {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));
typedef unsigned int u32x2v __attribute__((ext_vector_type(2)));
// four floats to bf16 (round to nearest even), two hardware conversions
__device__ inline bf16x4 bf16_4(float a, float b, float c, float d) {
  const bf16x2v lo = __builtin_convertvector((f32x2v){ a, b }, bf16x2v);
  const bf16x2v hi = __builtin_convertvector((f32x2v){ c, d }, bf16x2v);
  return __builtin_bit_cast(bf16x4, ((u32x2v){ __builtin_bit_cast(unsigned int, lo), __builtin_bit_cast(unsigned int, hi) }));
}
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {
constexpr int SG_NB = 32;
constexpr int SG_H = 128;
constexpr int SG_MAXT = 4;
constexpr int SG_ROWS = 64;   // GEMM rows per block
const float PI_F = 3.14159265358979324f;

__host__ __device__ inline unsigned short f2bf(float f) {
  unsigned int u;
  memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);   // round to nearest even
  return (unsigned short)(u >> 16);
}
__device__ inline float silu(float z) { return z / (1.0f + __expf(-z)); }
// logistic sigmoid with the hardware reciprocal (network inference)
__device__ inline float sigm_fast(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ inline float dsilu(float z) {
  float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}
}  // namespace

struct SurrogateParams {   // descriptor constants (by value to kernels)
  int nt, din;
  float rc, eta, mu0, dmu;
};

}  // namespace emdee

using namespace emdee;

struct EmDeeSurrogate {
  SurrogateParams sp;
  int hidden = SG_H;
  int P = 0;                                  // parameters per species
  std::vector<float> hostParams;              // [nt][P]
  DBuf<float> params, grads, am, av;          // fp32 master, gradient, Adam moments
  DBuf<unsigned short> w1t, w1, w2t, w2;      // bf16: [nt][H][din], [nt][din][H], [nt][H][H] x2
  DBuf<float> xmean, xstd;                    // descriptor normalisation
  std::vector<float> ymean, ystd;             // label normalisation per species
  bool normalised = false;
  long adamStep = 0;
  // stored training rows (raw descriptors, labels, species)
  std::vector<float> hostY;
  std::vector<int> hostS;
  DBuf<float> Xs, Ys;
  DBuf<int> Ss;
  long nrows = 0, caprows = 0;
  // training work buffers (rows padded to SG_ROWS)
  long workRows = 0;
  DBuf<float> X, Z1, H1, Z2, H2, G2, G1, dX, e, dXsum, part;
  DBuf<double> Fw, red;          // results; per-block partials
  // inference: compact list (rebuilt with the engine's list), normalised
  // descriptors, atom energies and dE/dD (slot order), partial sums
  DBuf<int> cJ, cN, cOver;
  int cap = 0;
  DBuf<int> rowSlot, grpSpecies;   // network rows grouped by species (k_sg_rows)
  int ngroups = 0;
  DBuf<float4> Rf;                 // single-precision positions and species
  DBuf<float4> nearE, nearG;       // near list of the latest descriptors (SG_NEAR per atom)
  DBuf<int> nearN, nearOver;
  long nearRows = 0;
  const void* listOwner = nullptr;   // Device whose list the compact list follows
  long listVersion = -1;
  long inferRows = 0;
  DBuf<float> Xi, dEdD, eat;
  DBuf<double> epart, wpart;
  // exact long-range channel: the reciprocal sum over |k| <= kcut (0: none)
  double kcut = 0.0;
  EwaldState* lowk = nullptr;
  bool queued = false, lowkQueued = false;   // dev_surrogate_launch ran; it queued the low-k sum
  const void* lowkOwner = nullptr;
  double lowkL = 0.0;
  DBuf<double> lowkAtom;         // low-k atom energies of a sample (slot order)
  hipStream_t own = nullptr;     // training stream
  hipStream_t stream = nullptr;  // stream of the current operation
};

namespace emdee {

// offsets inside one species' parameter block
struct ParamLayout {
  int din, H;
  __host__ __device__ int W1() const { return 0; }
  __host__ __device__ int b1() const { return din * H; }
  __host__ __device__ int W2() const { return din * H + H; }
  __host__ __device__ int b2() const { return din * H + H + H * H; }
  __host__ __device__ int w3() const { return din * H + 2 * H + H * H; }
  __host__ __device__ int b3() const { return din * H + 3 * H + H * H; }
  __host__ __device__ int size() const { return din * H + 3 * H + H * H + 1; }
};

// ---- descriptors ------------------------------------------------------------


// ---- inference: compact list, descriptors, fused MLP, forces ------------------
//
// The compact list of an atom holds its neighbours closer than rc + skin at
// the positions of the engine's latest list build (R0), taken from the
// engine's Verlet list in its order (sixteen lanes per atom, ordered by
// ballot).  It is rebuilt only when the engine's list is: under the engine's
// rebuild rule two atoms approach by at most the skin between builds, so
// every pair closer than rc at any later step is in it (the same guarantee
// the engine's own list relies on).
constexpr int SG_TPA = 16;

__global__ void __launch_bounds__(BLOCK) k_sg_list(int* __restrict__ cJ, int* __restrict__ cN, int* __restrict__ over,
                                                   const double* __restrict__ R0, const int* __restrict__ nbr,
                                                   const int* __restrict__ nnb, double rl2, double L, int N, int Npad,
                                                   int cap) {
  int gt = blockIdx.x * blockDim.x + threadIdx.x;
  int i = gt / SG_TPA, sub = gt % SG_TPA, lane = threadIdx.x & 63;
  bool valid = i < N;
  int n = valid ? nnb[i] : 0;
  const double invL = 1.0 / L;
  double xi = valid ? R0[i] : 0.0, yi = valid ? R0[Npad + i] : 0.0, zi = valid ? R0[2 * Npad + i] : 0.0;
  int base = 0;
  // n is uniform over the atom's sixteen lanes, so each ballot below sees
  // all of them
  for (int k0 = 0; k0 < n; k0 += SG_TPA) {
    int k = k0 + sub, j = 0;
    bool keep = false;
    if (k < n) {
      j = nbr[nb_at(k, i, Npad)];
      double dx = xi - R0[j], dy = yi - R0[Npad + j], dz = zi - R0[2 * Npad + j];
      dx -= L * rint(dx * invL);
      dy -= L * rint(dy * invL);
      dz -= L * rint(dz * invL);
      keep = dx * dx + dy * dy + dz * dz < rl2;
    }
    unsigned long long m = __ballot(keep);
    unsigned int gm = (unsigned int)(m >> (lane & ~(SG_TPA - 1))) & 0xffffu;
    int c = base + __popc(gm & ((1u << sub) - 1u));
    if (keep && c < cap) cJ[(size_t)i * cap + c] = j;
    base += __popc(gm);
  }
  if (valid && sub == 0) {
    cN[i] = min(base, cap);
    if (base > cap) atomicMax(over, base);
  }
}

// Rows of the network grouped by species (deterministic: ascending slot
// within a species), each species padded to whole 64-row groups; rowSlot = -1
// marks padding and grpSpecies[g] the species of group g (-1 past the end).
// One block of 1024 threads: a scan of the species flags per species.
__global__ void __launch_bounds__(1024) k_sg_rows(int* __restrict__ rowSlot, int* __restrict__ grpSpecies,
                                                  const int* __restrict__ type, int nt, int N, int ngroups) {
  __shared__ int part[1024];
  __shared__ int base;
  const int t = threadIdx.x, chunk = (N + 1023) / 1024;
  const int a0 = min(N, t * chunk), a1 = min(N, a0 + chunk);
  for (int g = t; g < ngroups; g += 1024) grpSpecies[g] = -1;
  if (t == 0) base = 0;
  __syncthreads();
  for (int s = 0; s < nt; s++) {
    int c = 0;
    for (int a = a0; a < a1; a++) c += type[a] == s;
    part[t] = c;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {   // inclusive scan (Hillis-Steele)
      int v = t >= off ? part[t - off] : 0;
      __syncthreads();
      part[t] += v;
      __syncthreads();
    }
    int pos = base + part[t] - c;
    for (int a = a0; a < a1; a++)
      if (type[a] == s) rowSlot[pos++] = a;
    const int cnt = part[1023], padded = (cnt + 63) & ~63;
    for (int r = base + cnt + t; r < base + padded; r += 1024) rowSlot[r] = -1;
    for (int g = base / 64 + t; g < (base + padded) / 64; g += 1024) grpSpecies[g] = s;
    __syncthreads();
    if (t == 0) base += padded;
    __syncthreads();
  }
}

// Single-precision positions and species for the gathers of the descriptor
// kernel: one 16-byte load per neighbour.
__global__ void k_sg_pos(float4* __restrict__ Rf, const double* __restrict__ R, const int* __restrict__ type, int N,
                         int Npad) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) Rf[i] = make_float4((float)R[i], (float)R[Npad + i], (float)R[2 * Npad + i], __int_as_float(type[i]));
}

__device__ __forceinline__ float sg_wrap(float d, float L, float invL) { return d - L * rintf(d * invL); }

// Near list: the entries of an atom's compact list closer than rc at the
// current positions, at most SG_NEAR per atom: {r_ij / r, r} and
// {fc, dfc, j, t_j} (cutoff function and its derivative, neighbour slot and
// species), written by the descriptor kernel for the force kernel.
constexpr int SG_NEAR = 256;
constexpr int SG_G = 32;   // lanes per Gaussian set: lane g of a stream owns Gaussian g

// Descriptors, one wavefront per atom.  Phase A: the 64 lanes test the
// atom's compact-list entries (four consecutive indices per lane in one load,
// then four position gathers in flight together) and compact those within
// rc, in a fixed order (ballot prefix), into LDS and into the near list.  Each LDS record holds r and the
// cutoff factor once per neighbour species (zero for the other species), and
// the records are padded with zeros to a multiple of eight.  Phase B: two
// streams of 32 lanes take the even and the odd records; lane g of a stream
// owns Gaussian g and adds it, times each species' factor, to that species'
// sum (broadcast LDS reads; every evaluated Gaussian is of an in-range pair
// or a zero-weight pad); the two streams' sums are added in a fixed order.
// Features are written normalised (mean != null) or raw.
template <int NT>
__global__ void __launch_bounds__(BLOCK) k_sg_desc(float* __restrict__ X, float4* __restrict__ nearE,
                                                   float4* __restrict__ nearG, int* __restrict__ nearN,
                                                   int* __restrict__ over, const int* __restrict__ cJ,
                                                   const int* __restrict__ cN, const float4* __restrict__ Rf,
                                                   const float* __restrict__ mean, const float* __restrict__ sd,
                                                   SurrogateParams sp, float L, int N, int cap) {
  constexpr int DIN = NT * SG_NB, AW = BLOCK / 64, SR = NT < 4 ? 4 : 8, SQ = SR / 4;
  __shared__ float4 ent[AW][SG_NEAR * SQ];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i = blockIdx.x * AW + wv;
  if (i >= N) return;   // uniform over the wave; waves share no LDS and no barrier
  float4* rec = ent[wv];
  const float invL = 1.0f / L, rc2 = sp.rc * sp.rc, kpi = PI_F / sp.rc, k2 = -sp.eta * 1.4426950408889634f;
  const int n = cN[i];
  const float4 ri = Rf[i];
  const int* __restrict__ row = cJ + (size_t)i * cap;
  auto put = [&](int pos, float r, float fc, int tj) {
    float v[SR];
#pragma unroll
    for (int k = 0; k < SR; k++) v[k] = 0.0f;
    v[0] = r;
#pragma unroll
    for (int t = 0; t < NT; t++) v[1 + t] = (t == tj) ? fc : 0.0f;
#pragma unroll
    for (int q = 0; q < SQ; q++) rec[pos * SQ + q] = make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
  };
  int cnt = 0;
  for (int c0 = 0; c0 < n; c0 += 64 * 4) {
    int j[4];
    float4 rj[4];
    {
      // lane takes entries cb..cb+3 with one 16-byte load (rows hold a
      // multiple of eight entries); a lane past the row's end reads its last
      // four, all beyond n
      const int cb = c0 + 4 * lane;
      const int4 jv = *reinterpret_cast<const int4*>(row + min(cb, cap - 4));
      j[0] = cb < n ? jv.x : -1;
      j[1] = cb + 1 < n ? jv.y : -1;
      j[2] = cb + 2 < n ? jv.z : -1;
      j[3] = cb + 3 < n ? jv.w : -1;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) rj[u] = Rf[j[u] >= 0 ? j[u] : i];
    // all four gathers are issued before the first is used
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < 4; u++) {
      float dx = sg_wrap(ri.x - rj[u].x, L, invL), dy = sg_wrap(ri.y - rj[u].y, L, invL),
            dz = sg_wrap(ri.z - rj[u].z, L, invL);
      float r2 = dx * dx + dy * dy + dz * dz;
      bool in = j[u] >= 0 && r2 < rc2;
      unsigned long long m = __ballot(in);
      int pos = cnt + __popcll(m & ((1ULL << lane) - 1ULL));
      if (in && pos < SG_NEAR) {
        float ir = __builtin_amdgcn_rsqf(r2), r = r2 * ir;
        float fc = 0.5f * (__cosf(kpi * r) + 1.0f), dfc = -0.5f * kpi * __sinf(kpi * r);
        int tj = __float_as_int(rj[u].w);
        put(pos, r, fc, tj);
        nearE[(size_t)i * SG_NEAR + pos] = make_float4(dx * ir, dy * ir, dz * ir, r);
        nearG[(size_t)i * SG_NEAR + pos] = make_float4(fc, dfc, __int_as_float(j[u]), __int_as_float(tj));
      }
      cnt += __popcll(m);
    }
  }
  if (lane == 0) {
    nearN[i] = min(cnt, SG_NEAR);
    if (cnt > SG_NEAR) atomicMax(over, cnt);
  }
  cnt = min(cnt, SG_NEAR);
  const int cpad = (cnt + 7) & ~7;
  if (cnt + lane < cpad) put(cnt + lane, 0.0f, 0.0f, -1);
  // the records are read by other lanes of this wave (LDS operations of one
  // wave complete in order)
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  const int g = lane & (SG_G - 1), st = lane >> 5;
  float acc[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) acc[t] = 0.0f;
  const float mu = sp.mu0 + g * sp.dmu;
  for (int e0 = 0; e0 < cpad; e0 += 8) {
    float v[4][SR];
#pragma unroll
    for (int u = 0; u < 4; u++) {
#pragma unroll
      for (int q = 0; q < SQ; q++) {
        const float4 w = rec[(e0 + 2 * u + st) * SQ + q];
        v[u][4 * q] = w.x; v[u][4 * q + 1] = w.y; v[u][4 * q + 2] = w.z; v[u][4 * q + 3] = w.w;
      }
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      float du = v[u][0] - mu, gv = __builtin_amdgcn_exp2f(k2 * du * du);
#pragma unroll
      for (int t = 0; t < NT; t++) acc[t] += gv * v[u][1 + t];
    }
  }
#pragma unroll
  for (int t = 0; t < NT; t++) acc[t] += __shfl_xor(acc[t], SG_G, 64);
  if (st != 0) return;
#pragma unroll
  for (int t = 0; t < NT; t++) {
    int f = t * SG_NB + g;
    X[(size_t)i * DIN + f] = mean ? (acc[t] - mean[f]) / sd[f] : acc[t];
  }
}

struct SgNorm { float ymean[SG_MAXT], ystd[SG_MAXT]; };

// The network of four 16-atom tiles of one species per workgroup (one tile
// per wavefront), forward and backward.  The species' bf16 weights are staged
// once in LDS (rows padded by four values, so the sixteen rows a fragment
// load touches fall on distinct banks).  Activations are kept transposed
// (features x atoms): the output fragment of v_mfma_f32_16x16x16_bf16 (lane
// holds features 4 (lane/16) + r of atom lane % 16) is exactly the B operand
// of the next product, so no activation leaves the registers.  Outputs: atom
// energies, dE/dD of the raw descriptors (slot order) and the group's energy
// sum (waves added in a fixed order).
template <int DIN>
struct SgLds {
  static constexpr int S1 = DIN + 4, SH = SG_H + 4;
  static constexpr int W1T = 0, W2T = W1T + SG_H * S1, W2 = W2T + SG_H * SH, W1 = W2 + SG_H * SH, SIZE = W1 + DIN * SH;
};
template <int NT>
__global__ void __launch_bounds__(256) k_sg_mlp(float* __restrict__ eat, float* __restrict__ dEdD,
                                                double* __restrict__ epart, const float* __restrict__ X,
                                                const int* __restrict__ rowSlot, const int* __restrict__ grpSpecies,
                                                const unsigned short* __restrict__ w1t,
                                                const unsigned short* __restrict__ w2t,
                                                const unsigned short* __restrict__ w2,
                                                const unsigned short* __restrict__ w1,
                                                const float* __restrict__ params, int P,
                                                const float* __restrict__ xstd, SgNorm nrm) {
  constexpr int DIN = NT * SG_NB, KT = DIN / 16, HT = SG_H / 16;
  using Lay = SgLds<DIN>;
  __shared__ unsigned short lw[Lay::SIZE];
  __shared__ float lb[3 * SG_H];   // b1, b2, w3
  __shared__ double wsum[4];
  const int s = grpSpecies[blockIdx.x];
  if (s < 0) {   // uniform over the block
    if (threadIdx.x == 0) epart[blockIdx.x] = 0.0;
    return;
  }
  // Loads in the order of use, since a wait for one load also waits for
  // every older one: W1^T (16-byte chunks), the biases b1, b2 and output
  // weights w3, the tile's slots and features, then W2^T, W2 and W1.  Each
  // matrix is written into its padded LDS rows just before the product that
  // reads it, so the later matrices arrive meanwhile.
  constexpr int C1 = SG_H * DIN / 8 / 256, C2 = SG_H * SG_H / 8 / 256;
  static_assert(C1 * 8 * 256 == SG_H * DIN && C2 * 8 * 256 == SG_H * SG_H, "whole chunks per thread");
  static_assert(SG_H <= 256, "one bias per thread");
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, col = lane & 15, g = lane >> 4;
  const ParamLayout pl{ DIN, SG_H };
  const float* p = params + (size_t)s * P;
  uint4 cw1t[C1], cw2t[C2], cw2[C2], cw1[C1];
  float bv[3];
  const uint4* s1t = reinterpret_cast<const uint4*>(w1t + (size_t)s * DIN * SG_H);
#pragma unroll
  for (int u = 0; u < C1; u++) cw1t[u] = s1t[u * 256 + threadIdx.x];
  if (threadIdx.x < SG_H) {
    bv[0] = p[pl.b1() + threadIdx.x];
    bv[1] = p[pl.b2() + threadIdx.x];
    bv[2] = p[pl.w3() + threadIdx.x];
  }
  const int slot = rowSlot[blockIdx.x * 64 + wv * 16 + col];
  float4 xv[KT];
#pragma unroll
  for (int q = 0; q < KT; q++)   // unconditional (slot 0 for an empty row), zeroed when converted
    xv[q] = *reinterpret_cast<const float4*>(X + (size_t)max(slot, 0) * DIN + 16 * q + 4 * g);
  __builtin_amdgcn_sched_barrier(0);   // keep the order of the loads
  {
    const uint4* s2t = reinterpret_cast<const uint4*>(w2t + (size_t)s * SG_H * SG_H);
    const uint4* s2 = reinterpret_cast<const uint4*>(w2 + (size_t)s * SG_H * SG_H);
    const uint4* s1 = reinterpret_cast<const uint4*>(w1 + (size_t)s * DIN * SG_H);
#pragma unroll
    for (int u = 0; u < C2; u++) cw2t[u] = s2t[u * 256 + threadIdx.x];
#pragma unroll
    for (int u = 0; u < C2; u++) cw2[u] = s2[u * 256 + threadIdx.x];
#pragma unroll
    for (int u = 0; u < C1; u++) cw1[u] = s1[u * 256 + threadIdx.x];
  }
  // chunk c of a matrix with `cols` columns: row c / (cols / 8), column 8 (c % (cols / 8))
  auto put = [&](const uint4* v, int nchunk, int cols, int dst, int stride) {
    const int per = cols / 8;
#pragma unroll
    for (int u = 0; u < nchunk; u++) {
      const int c = u * 256 + threadIdx.x, r = c / per, k = 8 * (c % per);
      *reinterpret_cast<uint2*>(lw + dst + r * stride + k) = make_uint2(v[u].x, v[u].y);
      *reinterpret_cast<uint2*>(lw + dst + r * stride + k + 4) = make_uint2(v[u].z, v[u].w);
    }
  };
  put(cw1t, C1, DIN, Lay::W1T, Lay::S1);
  if (threadIdx.x < SG_H) {
    lb[threadIdx.x] = bv[0];
    lb[SG_H + threadIdx.x] = bv[1];
    lb[2 * SG_H + threadIdx.x] = bv[2];
  }
  __syncthreads();
  const float* b1 = lb;
  const float* b2 = lb + SG_H;
  const float* w3 = lb + 2 * SG_H;
  bf16x4 xb[KT];
#pragma unroll
  for (int q = 0; q < KT; q++) {
    const float4 v = slot >= 0 ? xv[q] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    xb[q] = bf16_4(v.x, v.y, v.z, v.w);
  }
  // d1, d2: SiLU'(Z1), SiLU'(Z2), formed with the forward activations from
  // one sigmoid per value
  f32x4 d1[HT], d2[HT];
  bf16x4 hb[HT], gb[HT];
  float part = 0.0f;
#pragma unroll
  for (int m = 0; m < HT; m++) {
    f32x4 acc = f32x4{ 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
    for (int q = 0; q < KT; q++) {
      bf16x4 a = *reinterpret_cast<const bf16x4*>(lw + Lay::W1T + (16 * m + col) * Lay::S1 + 16 * q + 4 * g);
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, xb[q], acc, 0, 0, 0);
    }
    float h[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const float z = acc[r] + b1[16 * m + 4 * g + r], sg = sigm_fast(z);
      h[r] = z * sg;
      d1[m][r] = sg * (1.0f + z * (1.0f - sg));
    }
    hb[m] = bf16_4(h[0], h[1], h[2], h[3]);
  }
  put(cw2t, C2, SG_H, Lay::W2T, Lay::SH);
  __syncthreads();
#pragma unroll
  for (int m = 0; m < HT; m++) {
    f32x4 acc = f32x4{ 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
    for (int q = 0; q < HT; q++) {
      bf16x4 a = *reinterpret_cast<const bf16x4*>(lw + Lay::W2T + (16 * m + col) * Lay::SH + 16 * q + 4 * g);
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, hb[q], acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const float z = acc[r] + b2[16 * m + 4 * g + r], sg = sigm_fast(z);
      part += w3[16 * m + 4 * g + r] * (z * sg);
      d2[m][r] = sg * (1.0f + z * (1.0f - sg));
    }
  }
  part += __shfl_xor(part, 16, 64);
  part += __shfl_xor(part, 32, 64);
  const float e = nrm.ystd[s] * (part + p[pl.b3()]) + nrm.ymean[s];
  const float seed = slot >= 0 ? nrm.ystd[s] : 0.0f;
#pragma unroll
  for (int m = 0; m < HT; m++) {
    const float* w = w3 + 16 * m + 4 * g;
    gb[m] = bf16_4(seed * w[0] * d2[m][0], seed * w[1] * d2[m][1], seed * w[2] * d2[m][2], seed * w[3] * d2[m][3]);
  }
  put(cw2, C2, SG_H, Lay::W2, Lay::SH);
  put(cw1, C1, SG_H, Lay::W1, Lay::SH);
  __syncthreads();
  // G1 = (W2 G2) * SiLU'(Z1), into hb
#pragma unroll
  for (int m = 0; m < HT; m++) {
    f32x4 acc = f32x4{ 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
    for (int q = 0; q < HT; q++) {
      bf16x4 a = *reinterpret_cast<const bf16x4*>(lw + Lay::W2 + (16 * m + col) * Lay::SH + 16 * q + 4 * g);
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, gb[q], acc, 0, 0, 0);
    }
    hb[m] = bf16_4(acc[0] * d1[m][0], acc[1] * d1[m][1], acc[2] * d1[m][2], acc[3] * d1[m][3]);
  }
  // dE/dX = W1 G1
  f32x4 dx[KT];
#pragma unroll
  for (int m = 0; m < KT; m++) {
    dx[m] = f32x4{ 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
    for (int q = 0; q < HT; q++) {
      bf16x4 a = *reinterpret_cast<const bf16x4*>(lw + Lay::W1 + (16 * m + col) * Lay::SH + 16 * q + 4 * g);
      dx[m] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, hb[q], dx[m], 0, 0, 0);
    }
  }
  if (slot >= 0) {
#pragma unroll
    for (int m = 0; m < KT; m++) {
      int f = 16 * m + 4 * g;
      const float4 sd = *reinterpret_cast<const float4*>(xstd + f);
      float4 v = make_float4(dx[m][0] * __builtin_amdgcn_rcpf(sd.x), dx[m][1] * __builtin_amdgcn_rcpf(sd.y),
                             dx[m][2] * __builtin_amdgcn_rcpf(sd.z), dx[m][3] * __builtin_amdgcn_rcpf(sd.w));
      *reinterpret_cast<float4*>(dEdD + (size_t)slot * DIN + f) = v;
    }
    if (g == 0) eat[slot] = e;
  }
  double v = (g == 0 && slot >= 0) ? (double)e : 0.0;
  v = wave_sum(v);
  if (lane == 0) wsum[wv] = v;
  __syncthreads();
  if (threadIdx.x == 0) epart[blockIdx.x] = (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
}

// Forces and virial from dE/dD over the near list, one wavefront per atom:
// two streams of 32 lanes take the even and the odd near entries; lane g of
// a stream owns Gaussian g and sums its share of
// dE/dr_ij = sum_g (dE/dD_i[t_j][g] + dE/dD_j[t_i][g]) dG_g/dr (the pair
// moves both atoms' descriptors) times -r_ij / r.  The wave first copies a
// chunk of up to 128 near entries into LDS with coalesced loads (one entry
// per lane), padded to a multiple of eight with entries of zero cutoff
// factors (j = i), which contribute exactly zero; the streams then read the
// entries by broadcast from LDS and gather only dE/dD_j.  The 64 lanes' sums
// are combined once per atom by a fixed butterfly.  The list is full, so
// each atom sums its own force (no atomics).  Virial: -1/2 dE/dr r per entry.
// The near list carries the unit vector r_ij / r, so no division per pair.
template <int NT>
__global__ void __launch_bounds__(BLOCK) k_sg_force(double* __restrict__ F, double* __restrict__ wpart,
                                                    const float* __restrict__ dEdD, const float4* __restrict__ nearE,
                                                    const float4* __restrict__ nearG, const int* __restrict__ nearN,
                                                    const int* __restrict__ type, SurrogateParams sp, int N,
                                                    int Npad) {
  constexpr int DIN = NT * SG_NB, AW = BLOCK / 64, CH = 128;
  __shared__ float4 stE[AW][CH], stG[AW][CH];
  __shared__ double wred[AW];
  const int lane = threadIdx.x & 63, g = lane & (SG_G - 1), st = lane >> 5;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i = blockIdx.x * AW + wv;
  float fx = 0.0f, fy = 0.0f, fz = 0.0f, w = 0.0f;
  if (i < N) {
    const float k2 = -sp.eta * 1.4426950408889634f, mu = sp.mu0 + g * sp.dmu, eta2 = 2.0f * sp.eta;
    const int ti = type[i], cnt = nearN[i];
    float ai[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) ai[t] = dEdD[(size_t)i * DIN + t * SG_NB + g];
    const float4* E = nearE + (size_t)i * SG_NEAR;
    const float4* G = nearG + (size_t)i * SG_NEAR;
    const float4 padE = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float4 padG = make_float4(0.0f, 0.0f, __int_as_float(i), __int_as_float(ti));
    // byte offset of dE/dD_j[t_i][g] is 4 (j DIN + t_i SG_NB + g)
    const char* dj = reinterpret_cast<const char*>(dEdD);
    const unsigned int djo = 4u * (unsigned int)(ti * SG_NB + g);
    float4* sE = stE[wv];
    float4* sG = stG[wv];
    for (int c0 = 0; c0 < cnt; c0 += CH) {
      const int len = min(cnt - c0, CH);
      // the previous chunk's reads are done before it is overwritten (LDS
      // operations of one wave complete in order)
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int h = 0; h < CH / 64; h++) {
        const int e = h * 64 + lane, ec = c0 + min(e, len - 1);
        float4 ve = E[ec], vg = G[ec];
        if (e >= len) { ve = padE; vg = padG; }
        sE[e] = ve;
        sG[e] = vg;
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      for (int k0 = 0; k0 < len; k0 += 8) {
        float4 ec[4], gc[4];
        float aj[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          ec[u] = sE[k0 + 2 * u + st];
          gc[u] = sG[k0 + 2 * u + st];
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
          aj[u] = *reinterpret_cast<const float*>(dj + (4u * DIN * (unsigned int)__float_as_int(gc[u].z) + djo));
#pragma unroll
        for (int u = 0; u < 4; u++) {
          int tj = __float_as_int(gc[u].w);
          float sel = ai[0];
#pragma unroll
          for (int t = 1; t < NT; t++) sel = (tj == t) ? ai[t] : sel;
          float du = ec[u].w - mu, gauss = __builtin_amdgcn_exp2f(k2 * du * du);
          float dEdr = (sel + aj[u]) * (gauss * (gc[u].y - eta2 * du * gc[u].x));
          fx -= dEdr * ec[u].x;
          fy -= dEdr * ec[u].y;
          fz -= dEdr * ec[u].z;
          w -= dEdr * ec[u].w;
        }
      }
    }
  }
  double dx = fx, dy = fy, dz = fz, dw = 0.5 * (double)w;
  dx = wave_sum(dx);
  dy = wave_sum(dy);
  dz = wave_sum(dz);
  dw = wave_sum(dw);
  if (lane == 0) {
    if (i < N) {
      F[i] += dx;
      F[Npad + i] += dy;
      F[2 * Npad + i] += dz;
    }
    wred[wv] = dw;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s2 = 0.0;
    for (int k = 0; k < AW; k++) s2 += wred[k];
    wpart[blockIdx.x] = s2;
  }
}

__global__ void k_sg_normalise(float* __restrict__ X, const float* __restrict__ D, const float* __restrict__ mean,
                               const float* __restrict__ sd, long rows, int din, long padRows) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= padRows * din) return;
  long r = idx / din;
  int f = (int)(idx % din);
  X[idx] = r < rows ? (D[idx] - mean[f]) / sd[f] : 0.0f;
}

// ---- MLP on MFMA --------------------------------------------------------------
// Z[M x NC] = A[M x K] . B[K x NC] (+ bias); H = SiLU(Z) when ACT.  Bt is B
// transposed, [NC][K] in bf16, so a lane loads its 4 K-values contiguously.
// Lane layout of v_mfma_f32_16x16x16_bf16: A[i][k] with i = lane%16,
// k = 4 (lane/16) + r; B[k][j] with j = lane%16; D[i][j] with i = 4 (lane/16) + r.
enum { EPI_Z = 0, EPI_SILU = 1, EPI_DSILU = 2, EPI_ADD = 3 };
// One block of 4 wavefronts computes 16 rows x NC columns; wave w takes the
// column tiles w, w+4, ...  Epilogues: EPI_Z stores Z = A.B (+ bias);
// EPI_SILU also stores H = SiLU(Z); EPI_DSILU stores Z * SiLU'(aux);
// EPI_ADD adds Z into the output.
template <int K, int NC, int EPI>
__global__ void __launch_bounds__(256) k_sg_gemm(const float* __restrict__ A, const unsigned short* __restrict__ Bt,
                                                 const float* __restrict__ bias, float* __restrict__ Z,
                                                 float* __restrict__ Hout, const float* __restrict__ aux) {
  constexpr int TILES = NC / 16, PER = (TILES + 3) / 4;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const long r0 = (long)blockIdx.x * 16;
  const int ai = lane & 15, ak = 4 * (lane >> 4);
  f32x4 acc[PER];
#pragma unroll
  for (int m = 0; m < PER; m++) acc[m] = f32x4{ 0.0f, 0.0f, 0.0f, 0.0f };
  // all fragments are loaded before the MFMA chain so that the loads overlap
  float4 av[K / 16];
  bf16x4 bv[PER][K / 16];
#pragma unroll
  for (int q = 0; q < K / 16; q++) av[q] = *reinterpret_cast<const float4*>(A + (r0 + ai) * K + 16 * q + ak);
#pragma unroll
  for (int m = 0; m < PER; m++) {
    int t = w + 4 * m < TILES ? w + 4 * m : TILES - 1;
#pragma unroll
    for (int q = 0; q < K / 16; q++)
      bv[m][q] = *reinterpret_cast<const bf16x4*>(Bt + (size_t)(16 * t + ai) * K + 16 * q + ak);
  }
#pragma unroll
  for (int q = 0; q < K / 16; q++) {
    bf16x4 a = { (short)f2bf(av[q].x), (short)f2bf(av[q].y), (short)f2bf(av[q].z), (short)f2bf(av[q].w) };
#pragma unroll
    for (int m = 0; m < PER; m++)
      if (w + 4 * m < TILES) acc[m] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bv[m][q], acc[m], 0, 0, 0);
  }
#pragma unroll
  for (int m = 0; m < PER; m++) {
    int t = w + 4 * m;
    if (t >= TILES) continue;
    int j = 16 * t + ai;
    float bj = bias ? bias[j] : 0.0f;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      long o = (r0 + ak + r) * NC + j;
      float z = acc[m][r] + bj;
      if (EPI == EPI_Z) Z[o] = z;
      else if (EPI == EPI_SILU) { Z[o] = z; Hout[o] = silu(z); }
      else if (EPI == EPI_DSILU) Z[o] = z * dsilu(aux[o]);
      else Z[o] += z;
    }
  }
}

// Weight gradient partials: part[block][K x NC] = sum over the block's rows of
// A[row][k] G[row][n], on MFMA with the rows as the reduction dimension.
template <int K, int NC>
__global__ void __launch_bounds__(256) k_sg_wgrad(float* __restrict__ part, const float* __restrict__ A,
                                                  const float* __restrict__ G, long rows, int rowsPerBlock) {
  constexpr int TK = K / 16, TN = NC / 16, TILES = TK * TN, PER = TILES / 4;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, lk = 4 * (lane >> 4);
  f32x4 acc[PER];
#pragma unroll
  for (int m = 0; m < PER; m++) acc[m] = f32x4{ 0.0f, 0.0f, 0.0f, 0.0f };
  long rb = (long)blockIdx.x * rowsPerBlock, re = rb + rowsPerBlock < rows ? rb + rowsPerBlock : rows;
  for (long c = rb; c < re; c += 16) {
#pragma unroll
    for (int m = 0; m < PER; m++) {
      int t = w + 4 * m, ti = t / TN, tj = t % TN;
      bf16x4 a, b;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        long row = c + lk + r;
        bool ok = row < re;
        a[r] = (short)f2bf(ok ? A[row * K + 16 * ti + li] : 0.0f);
        b[r] = (short)f2bf(ok ? G[row * NC + 16 * tj + li] : 0.0f);
      }
      acc[m] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[m], 0, 0, 0);
    }
  }
  float* out = part + (size_t)blockIdx.x * K * NC;
#pragma unroll
  for (int m = 0; m < PER; m++) {
    int t = w + 4 * m, ti = t / TN, tj = t % TN;
#pragma unroll
    for (int r = 0; r < 4; r++) out[(16 * ti + lk + r) * NC + 16 * tj + li] = acc[m][r];
  }
}

// column sums part[block][n] = sum_rows G[row][n] (bias gradients), and
// optionally sum_rows g[row] H[row][n] (output-layer weights).
__global__ void k_sg_colsum(float* __restrict__ part, const float* __restrict__ G, const float* __restrict__ g,
                            int NC, long rows, int rowsPerBlock) {
  int n = threadIdx.x;
  if (n >= NC) return;
  long rb = (long)blockIdx.x * rowsPerBlock, re = rb + rowsPerBlock < rows ? rb + rowsPerBlock : rows;
  float s = 0.0f;
  for (long r = rb; r < re; r++) s += g ? g[r] * G[r * NC + n] : G[r * NC + n];
  part[(size_t)blockIdx.x * NC + n] = s;
}

// dst[n] = sum over blocks of part[block][n] (fixed order)
__global__ void k_sg_reduce(float* __restrict__ dst, const float* __restrict__ part, int nblocks, int n) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  float s = 0.0f;
  for (int b = 0; b < nblocks; b++) s += part[(size_t)b * n + idx];
  dst[idx] = s;
}

// Output layer of species s and the seed of the backward pass, one
// wavefront per row (two columns per lane).
//   mode 0 (inference): eatom[row] = ystd out + ymean for rows of species s,
//       seed = ystd (dE/d out).
//   mode 1 (training): loss partials of (out - y)^2, seed = 2 (out - y) / count.
// G2 = seed * w3 * SiLU'(Z2) (zero for rows of other species).
__global__ void __launch_bounds__(256) k_sg_output(float* __restrict__ G2, float* __restrict__ eatom,
                                                   float* __restrict__ seedOut, double* __restrict__ red,
                                                   const float* __restrict__ H2, const float* __restrict__ Z2,
                                                   const float* __restrict__ p, const int* __restrict__ species,
                                                   const float* __restrict__ y, int s, float ymean, float ystd,
                                                   float invCount, long rows, int mode, ParamLayout pl) {
  const int lane = threadIdx.x & 63;
  long row = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  double v[1] = { 0.0 };
  if (row < rows) {
    const float* w3 = p + pl.w3();
    const float* h = H2 + row * SG_H;
    float part = h[lane] * w3[lane] + h[lane + 64] * w3[lane + 64];
    float out = wave_sum(part) + p[pl.b3()];
    out = __shfl(out, 0, 64);
    bool mine = species[row] == s;
    float seed = 0.0f;
    if (mine) {
      if (mode == 0) {
        seed = ystd;
        if (lane == 0) { eatom[row] = ystd * out + ymean; v[0] = (double)(ystd * out + ymean); }
      } else {
        float dd = out - y[row];
        seed = 2.0f * dd * invCount;
        if (lane == 0) v[0] = (double)dd * (double)dd;
      }
    }
    if (lane == 0) seedOut[row] = seed;
    const float* z = Z2 + row * SG_H;
    float* g = G2 + row * SG_H;
    g[lane] = seed * w3[lane] * dsilu(z[lane]);
    g[lane + 64] = seed * w3[lane + 64] * dsilu(z[lane + 64]);
  }
  block_sum<1>(v);
  if (threadIdx.x == 0) red[blockIdx.x] = v[0];
}

// Adam (Kingma & Ba 2015) on the fp32 master parameters of one species.
__global__ void k_sg_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                          const float* __restrict__ g, int n, float lr, float b1, float b2, float eps, float c1,
                          float c2) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  float gi = g[idx];
  m[idx] = b1 * m[idx] + (1.0f - b1) * gi;
  v[idx] = b2 * v[idx] + (1.0f - b2) * gi * gi;
  p[idx] -= lr * (m[idx] / c1) / (sqrtf(v[idx] / c2) + eps);
}

// bf16 copies of W1 and W2 in both orientations
__global__ void k_sg_pack(unsigned short* __restrict__ w1t, unsigned short* __restrict__ w1,
                          unsigned short* __restrict__ w2t, unsigned short* __restrict__ w2,
                          const float* __restrict__ p, ParamLayout pl) {
  int din = pl.din, H = pl.H;
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < din * H) {
    int k = idx / H, n = idx % H;
    float x = p[pl.W1() + idx];
    w1[idx] = f2bf(x);             // [din][H]: Bt of the product by W1^T
    w1t[n * din + k] = f2bf(x);    // [H][din]: Bt of the product by W1
  }
  if (idx < H * H) {
    int k = idx / H, n = idx % H;
    float x = p[pl.W2() + idx];
    w2[idx] = f2bf(x);
    w2t[n * H + k] = f2bf(x);
  }
}

// ---- host side -------------------------------------------------------------------

static void ensure_work(EmDeeSurrogate* s, long rows) {
  long pad = ((rows + SG_ROWS - 1) / SG_ROWS) * SG_ROWS;
  if (pad <= s->workRows) return;
  int din = s->sp.din, H = s->hidden;
  s->X.alloc(pad * din); s->dX.alloc(pad * din); s->dXsum.alloc(pad * din);
  for (DBuf<float>* b : { &s->Z1, &s->H1, &s->Z2, &s->H2, &s->G2, &s->G1 }) b->alloc(pad * H);
  s->e.alloc(pad * 2);
  s->red.alloc(pad / 4 + 64);    // >= one partial per block (four rows per block in k_sg_output)
  s->part.alloc((size_t)(pad / 256 + 2) * H * H);
  HIP_CHECK(hipMemsetAsync(s->X.p, 0, sizeof(float) * pad * din, s->stream));
  for (DBuf<float>* b : { &s->Z1, &s->H1, &s->Z2, &s->H2, &s->G2, &s->G1 })
    HIP_CHECK(hipMemsetAsync(b->p, 0, sizeof(float) * pad * H, s->stream));
  s->workRows = pad;
}

static void check_context(const EmDeeSurrogate* s, const Device* d, const SurrogateContext& c, const char* task) {
  if (s->sp.nt != d->ntypes) fatal(task, "the model and the system have different numbers of types");
  if ((double)s->sp.rc > c.Rc) fatal(task, "the descriptor cutoff exceeds the system's cutoff");
  if (!d->built) fatal(task, "the system has no neighbor list yet");
  // the force kernel gathers dE/dD with 32-bit byte offsets
  if ((double)d->Npad * s->sp.din * sizeof(float) >= 4294967296.0) fatal(task, "too many atoms for the model's kernels");
}

double surrogate_cutoff(const EmDeeSurrogate* s) { return (double)s->sp.rc; }

// The compact list of the system's current Verlet list: rebuilt when the
// engine has rebuilt its list since (or the model moved to another system).
static void ensure_list(EmDeeSurrogate* s, Device* d, const SurrogateContext& c) {
  if (s->listOwner == d && s->listVersion == d->listVersion) return;
  int N = d->N, Np = d->Npad;
  double rl = (double)s->sp.rc + c.skin, rl2 = rl * rl;
  if (s->cap == 0) {
    double vol = 4.18879020478639 * rl * rl * rl * ((double)N / (c.L * c.L * c.L));
    s->cap = ((int)(1.3 * vol + 32) + 7) & ~7;
  }
  s->cN.alloc(Np);
  s->cOver.alloc(1);
  for (;;) {
    s->cJ.alloc((size_t)s->cap * Np);
    HIP_CHECK(hipMemsetAsync(s->cOver.p, 0, sizeof(int), d->stream));
    hipLaunchKernelGGL(k_sg_list, dim3(div_up((long)Np * SG_TPA, BLOCK)), dim3(BLOCK), 0, d->stream, s->cJ.p, s->cN.p,
                       s->cOver.p, d->R0.p, d->nbr.p, d->nnb.p, rl2, c.L, N, Np, s->cap);
    launch_check("surrogate list");
    int over = 0;
    HIP_CHECK(hipMemcpyAsync(&over, s->cOver.p, sizeof(int), hipMemcpyDeviceToHost, d->stream));
    HIP_CHECK(hipStreamSynchronize(d->stream));
    if (over == 0) break;
    s->cap = ((int)(over * 1.2) + 15) & ~7;
  }
  // network rows by species, in the new slot order
  s->ngroups = d->Npad / 64 + s->sp.nt;
  s->rowSlot.alloc((size_t)64 * s->ngroups);
  s->grpSpecies.alloc(s->ngroups);
  hipLaunchKernelGGL(k_sg_rows, dim3(1), dim3(1024), 0, d->stream, s->rowSlot.p, s->grpSpecies.p, d->type.p, s->sp.nt,
                     N, s->ngroups);
  launch_check("surrogate rows");
  s->listOwner = d;
  s->listVersion = d->listVersion;
}

// Descriptors of the system's current positions into D ([N][din], rows in
// slot order), normalised or raw.
static void launch_desc(EmDeeSurrogate* s, Device* d, const SurrogateContext& c, float* D, bool normalised) {
  ensure_list(s, d, c);
  int Np = d->Npad;
  s->Rf.alloc(Np);
  if (s->nearRows < Np) {
    s->nearE.alloc((size_t)Np * SG_NEAR);
    s->nearG.alloc((size_t)Np * SG_NEAR);
    s->nearN.alloc(Np);
    s->nearOver.alloc(1);
    s->nearRows = Np;
  }
  HIP_CHECK(hipMemsetAsync(s->nearOver.p, 0, sizeof(int), d->stream));
  hipLaunchKernelGGL(k_sg_pos, dim3(div_up(d->N, BLOCK)), dim3(BLOCK), 0, d->stream, s->Rf.p, d->R.p, d->type.p, d->N,
                     Np);
  int g = div_up(d->N, BLOCK / 64);
  const float* m = normalised ? s->xmean.p : nullptr;
  const float* sd = normalised ? s->xstd.p : nullptr;
#define SG_DESC(NT)                                                                                                 \
  hipLaunchKernelGGL(k_sg_desc<NT>, dim3(g), dim3(BLOCK), 0, d->stream, D, s->nearE.p, s->nearG.p, s->nearN.p,      \
                     s->nearOver.p, s->cJ.p, s->cN.p, s->Rf.p, m, sd, s->sp, (float)c.L, d->N, s->cap)
  switch (s->sp.nt) { case 1: SG_DESC(1); break; case 2: SG_DESC(2); break; case 3: SG_DESC(3); break; default: SG_DESC(4); }
#undef SG_DESC
  launch_check("surrogate descriptors");
}

// The low-k state for this system and box (rebuilt when either changes).
static EwaldState* ensure_lowk(EmDeeSurrogate* s, Device* d, const SurrogateContext& c, const char* task) {
  if (s->kcut <= 0.0) return nullptr;
  if (s->lowk && s->lowkOwner == d && s->lowkL == c.L) return s->lowk;
  if (!c.ewald) fatal(task, "the long-range channel needs the system's kspace model");
  dev_lowk_release(s->lowk);
  s->lowk = dev_lowk_create(lowk_setup(*c.ewald, c.lambdaLayer, c.alpha, c.L, s->kcut));
  s->lowkOwner = d;
  s->lowkL = c.L;
  return s->lowk;
}

static void gemm(EmDeeSurrogate* s, int K, int NC, int epi, const float* A, const unsigned short* Bt, const float* bias,
                 float* Z, float* H, const float* aux, long padRows) {
  dim3 grid((unsigned)(padRows / 16));
#define SG_GEMM(KK, NN, EE) hipLaunchKernelGGL((k_sg_gemm<KK, NN, EE>), grid, dim3(256), 0, s->stream, A, Bt, bias, Z, H, aux)
#define SG_SHAPE(KK, NN)                                                                            \
  do {                                                                                              \
    if (epi == EPI_Z) SG_GEMM(KK, NN, EPI_Z); else if (epi == EPI_SILU) SG_GEMM(KK, NN, EPI_SILU); \
    else if (epi == EPI_DSILU) SG_GEMM(KK, NN, EPI_DSILU); else SG_GEMM(KK, NN, EPI_ADD);          \
  } while (0)
  if (K == 64 && NC == 128) SG_SHAPE(64, 128);
  else if (K == 128 && NC == 128) SG_SHAPE(128, 128);
  else if (K == 128 && NC == 64) SG_SHAPE(128, 64);
  else if (K == 32 && NC == 128) SG_SHAPE(32, 128);
  else if (K == 128 && NC == 32) SG_SHAPE(128, 32);
  else if (K == 96 && NC == 128) SG_SHAPE(96, 128);
  else if (K == 128 && NC == 96) SG_SHAPE(128, 96);
  else fatal("surrogate", "unsupported layer shape");
#undef SG_SHAPE
#undef SG_GEMM
  launch_check("surrogate gemm");
}

// Forward pass of species sp over all rows, then the backward pass to dX
// (added into dXsum) and, in training mode, the weight gradients.
static double mlp_pass(EmDeeSurrogate* s, int spc, long rows, const int* species, const float* y, float invCount,
                       int mode, float* eatom) {
  int din = s->sp.din, H = s->hidden;
  ParamLayout pl{ din, H };
  long pad = ((rows + SG_ROWS - 1) / SG_ROWS) * SG_ROWS;
  const float* p = s->params.p + (size_t)spc * s->P;
  const unsigned short* w1t = s->w1t.p + (size_t)spc * din * H;
  const unsigned short* w1 = s->w1.p + (size_t)spc * din * H;
  const unsigned short* w2t = s->w2t.p + (size_t)spc * H * H;
  const unsigned short* w2 = s->w2.p + (size_t)spc * H * H;
  gemm(s, din, H, EPI_SILU, s->X.p, w1t, p + pl.b1(), s->Z1.p, s->H1.p, nullptr, pad);
  gemm(s, H, H, EPI_SILU, s->H1.p, w2t, p + pl.b2(), s->Z2.p, s->H2.p, nullptr, pad);
  int gr = div_up(rows, 4);
  float* seed = s->e.p + pad;   // second half of e holds the per-row seeds
  hipLaunchKernelGGL(k_sg_output, dim3(gr), dim3(BLOCK), 0, s->stream, s->G2.p, eatom, seed, s->red.p, s->H2.p, s->Z2.p, p,
                     species, y, spc, s->ymean[spc], s->ystd[spc], invCount, rows, mode, pl);
  launch_check("surrogate output");
  // G1 = (G2 W2^T) * SiLU'(Z1)
  gemm(s, H, H, EPI_DSILU, s->G2.p, w2, nullptr, s->G1.p, nullptr, s->Z1.p, pad);
  if (mode == 0) {
    // dE/dX of this species' rows, added to the other species' (first pass stores)
    gemm(s, H, din, spc == 0 ? EPI_Z : EPI_ADD, s->G1.p, w1, nullptr, s->dXsum.p, nullptr, nullptr, pad);
  } else {
    // weight gradients of this species into grads
    float* g = s->grads.p + (size_t)spc * s->P;
    int rpb = 256, nb = (int)((rows + rpb - 1) / rpb);
#define SG_WGRAD(KK, NN, AP, GP, DST)                                                                          \
  do {                                                                                                        \
    hipLaunchKernelGGL((k_sg_wgrad<KK, NN>), dim3(nb), dim3(256), 0, s->stream, s->part.p, AP, GP, rows, rpb); \
    hipLaunchKernelGGL(k_sg_reduce, dim3(div_up(KK * NN, BLOCK)), dim3(BLOCK), 0, s->stream, DST, s->part.p, nb, \
                       KK * NN);                                                                              \
  } while (0)
    if (din == 64) SG_WGRAD(64, 128, s->X.p, s->G1.p, g + pl.W1());
    else if (din == 128) SG_WGRAD(128, 128, s->X.p, s->G1.p, g + pl.W1());
    else SG_WGRAD(32, 128, s->X.p, s->G1.p, g + pl.W1());
    SG_WGRAD(128, 128, s->H1.p, s->G2.p, g + pl.W2());
#undef SG_WGRAD
    auto colsum = [&](const float* G, const float* gw, float* dst) {
      hipLaunchKernelGGL(k_sg_colsum, dim3(nb), dim3(H), 0, s->stream, s->part.p, G, gw, H, rows, rpb);
      hipLaunchKernelGGL(k_sg_reduce, dim3(1), dim3(H), 0, s->stream, dst, s->part.p, nb, H);
    };
    colsum(s->G1.p, nullptr, g + pl.b1());
    colsum(s->G2.p, nullptr, g + pl.b2());
    colsum(s->H2.p, seed, g + pl.w3());
    // b3 gradient: sum of seeds
    hipLaunchKernelGGL(k_sg_colsum, dim3(nb), dim3(64), 0, s->stream, s->part.p, seed, nullptr, 1, rows, rpb);
    hipLaunchKernelGGL(k_sg_reduce, dim3(1), dim3(64), 0, s->stream, g + pl.b3(), s->part.p, nb, 1);
    launch_check("surrogate gradients");
  }
  // reduce the output-kernel partials (energy or loss)
  hipLaunchKernelGGL(k_reduce_partials<1>, dim3(1), dim3(BLOCK), 0, s->stream, s->Fw.p, s->red.p, gr);
  double h = 0;
  HIP_CHECK(hipMemcpyAsync(&h, s->Fw.p, sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIP_CHECK(hipStreamSynchronize(s->stream));
  return h;
}

static void repack(EmDeeSurrogate* s) {
  ParamLayout pl{ s->sp.din, s->hidden };
  for (int t = 0; t < s->sp.nt; t++) {
    int n = std::max(pl.din * pl.H, pl.H * pl.H);
    hipLaunchKernelGGL(k_sg_pack, dim3(div_up(n, BLOCK)), dim3(BLOCK), 0, s->stream,
                       s->w1t.p + (size_t)t * pl.din * pl.H, s->w1.p + (size_t)t * pl.din * pl.H,
                       s->w2t.p + (size_t)t * pl.H * pl.H, s->w2.p + (size_t)t * pl.H * pl.H,
                       s->params.p + (size_t)t * s->P, pl);
  }
  launch_check("surrogate pack");
}

// The model's energy and virial (fixed-order sums of the MLP group partials
// and the force kernel's block partials) and the near-list overflow count,
// posted into pinned host memory with sequence number seq.
__global__ void __launch_bounds__(BLOCK) k_sg_post(double* __restrict__ host, unsigned long long* __restrict__ hostSeq,
                                                   unsigned long long seq, const double* __restrict__ epart, int ne,
                                                   const double* __restrict__ wpart, int nw,
                                                   const int* __restrict__ over) {
  double v[2] = { 0.0, 0.0 };
  for (int b = threadIdx.x; b < ne; b += BLOCK) v[0] += epart[b];
  for (int b = threadIdx.x; b < nw; b += BLOCK) v[1] += wpart[b];
  block_sum<2>(v);
  if (threadIdx.x == 0) {
    host[0] = v[0];
    host[1] = v[1];
    host[2] = (double)*over;
    __threadfence_system();
    *(volatile unsigned long long*)hostSeq = seq;
  }
}

__global__ void k_sg_add_atom(double* __restrict__ dst, const float* __restrict__ src, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] += (double)src[i];
}

// Inference on a system's device state: adds the model forces to F(layer)
// (slot order) and its atom energies to the atom-energy array when that is
// on; returns the model energy (low-k sum plus network) and its virial in *W.
double dev_surrogate_apply(Device* d, EmDeeSurrogate* s, int layer, const SurrogateContext& c, double* W) {
  dev_surrogate_launch(d, s, layer, c);
  dev_wait_posts(d);
  return dev_surrogate_result(d, s, W);
}

double dev_surrogate_result(Device* d, EmDeeSurrogate* s, double* W) {
  *W = 0.0;
  if (!s->queued) return 0.0;
  s->queued = false;
  double El = 0.0, Wl = 0.0;
  if (s->lowkQueued) El = dev_lowk_result(d, &Wl);
  const double* hr = d->hpost + POST_MODEL;
  if (hr[2] != 0.0) fatal("surrogate", "more neighbours within the descriptor cutoff than the model holds (256)");
  *W = hr[1] + Wl;
  return hr[0] + El;
}

void dev_surrogate_launch(Device* d, EmDeeSurrogate* s, int layer, const SurrogateContext& c) {
  const char* task = "surrogate";
  if (!s->normalised) fatal(task, "the model has not been trained");
  int N = d->N, Np = d->Npad, din = s->sp.din;
  s->queued = false;
  s->lowkQueued = false;
  if (N == 0) return;
  check_context(s, d, c, task);
  s->stream = d->stream;
  double* F = d->F.p + 3 * (size_t)Np * layer;
  if (s->inferRows < Np) {
    s->Xi.alloc((size_t)Np * din);
    s->dEdD.alloc((size_t)Np * din);
    s->eat.alloc(Np);
    s->wpart.alloc(div_up(Np, BLOCK / 64));
    s->inferRows = Np;
  }
  launch_desc(s, d, c, s->Xi.p, true);
  s->epart.alloc(s->ngroups);
  SgNorm nrm;
  for (int t = 0; t < SG_MAXT; t++) {
    nrm.ymean[t] = t < s->sp.nt ? s->ymean[t] : 0.0f;
    nrm.ystd[t] = t < s->sp.nt ? s->ystd[t] : 0.0f;
  }
  int gf = div_up(N, BLOCK / 64);
#define SG_RUN(NT)                                                                                                     \
  do {                                                                                                                 \
    hipLaunchKernelGGL(k_sg_mlp<NT>, dim3(s->ngroups), dim3(256), 0, d->stream, s->eat.p, s->dEdD.p, s->epart.p,       \
                       s->Xi.p, s->rowSlot.p, s->grpSpecies.p, s->w1t.p, s->w2t.p, s->w2.p, s->w1.p, s->params.p, s->P, \
                       s->xstd.p, nrm);                                                                                \
    hipLaunchKernelGGL(k_sg_force<NT>, dim3(gf), dim3(BLOCK), 0, d->stream, F, s->wpart.p, s->dEdD.p, s->nearE.p,    \
                       s->nearG.p, s->nearN.p, d->type.p, s->sp, N, Np);                                               \
  } while (0)
  switch (s->sp.nt) { case 1: SG_RUN(1); break; case 2: SG_RUN(2); break; case 3: SG_RUN(3); break; default: SG_RUN(4); }
#undef SG_RUN
  launch_check("surrogate network and forces");
  const unsigned long long seq = ++d->seq;
  hipLaunchKernelGGL(k_sg_post, dim3(1), dim3(BLOCK), 0, d->stream, d->hpostDev + POST_MODEL, d->hseqDev, seq,
                     s->epart.p, s->ngroups, s->wpart.p, gf, s->nearOver.p);
  if (d->wantAtomE)
    hipLaunchKernelGGL(k_sg_add_atom, dim3(div_up(N, BLOCK)), dim3(BLOCK), 0, d->stream, d->eatom.p, s->eat.p, N);
  launch_check("surrogate sums");
  (void)seq;
  if (EwaldState* lk = ensure_lowk(s, d, c, task))
    s->lowkQueued = dev_lowk_launch(d, lk, F, d->wantAtomE ? d->eatom.p : nullptr, c.L);
  s->queued = true;
}

// Stores the raw descriptors of a system's current configuration as training
// rows, with per-atom labels given in original order; with the long-range
// channel the rows' labels are the residuals after its atom energies.
void dev_surrogate_add_sample(Device* d, EmDeeSurrogate* s, const SurrogateContext& c, const double* labels) {
  const char* task = "surrogate sampling";
  int N = d->N, din = s->sp.din;
  check_context(s, d, c, task);
  s->stream = d->stream;
  std::vector<int> iperm(N);
  HIP_CHECK(hipMemcpyAsync(iperm.data(), d->iperm.p, sizeof(int) * N, hipMemcpyDeviceToHost, d->stream));
  HIP_CHECK(hipStreamSynchronize(d->stream));
  std::vector<double> labelsSlot(N);
  for (int a = 0; a < N; a++) labelsSlot[iperm[a]] = labels[a];
  if (EwaldState* lk = ensure_lowk(s, d, c, task)) {
    // low-k atom energies at this configuration; its forces go to a scratch layer copy
    s->lowkAtom.alloc(d->Npad);
    HIP_CHECK(hipMemsetAsync(s->lowkAtom.p, 0, sizeof(double) * d->Npad, d->stream));
    DBuf<double> Fs;
    Fs.alloc(3 * (size_t)d->Npad);
    HIP_CHECK(hipMemsetAsync(Fs.p, 0, sizeof(double) * 3 * d->Npad, d->stream));
    double Wl = 0.0;
    dev_lowk_compute(d, lk, Fs.p, s->lowkAtom.p, c.L, &Wl);
    std::vector<double> el(N);
    HIP_CHECK(hipMemcpyAsync(el.data(), s->lowkAtom.p, sizeof(double) * N, hipMemcpyDeviceToHost, d->stream));
    HIP_CHECK(hipStreamSynchronize(d->stream));
    Fs.release();
    for (int i = 0; i < N; i++) labelsSlot[i] -= el[i];
  }
  if (s->nrows + N > s->caprows) {
    long cap = std::max(2 * s->caprows, s->nrows + N);
    DBuf<float> nx, ny;
    DBuf<int> ns;
    nx.alloc(cap * din); ny.alloc(cap); ns.alloc(cap);
    if (s->nrows) {
      HIP_CHECK(hipMemcpy(nx.p, s->Xs.p, sizeof(float) * s->nrows * din, hipMemcpyDeviceToDevice));
      HIP_CHECK(hipMemcpy(ny.p, s->Ys.p, sizeof(float) * s->nrows, hipMemcpyDeviceToDevice));
      HIP_CHECK(hipMemcpy(ns.p, s->Ss.p, sizeof(int) * s->nrows, hipMemcpyDeviceToDevice));
    }
    s->Xs.release(); s->Ys.release(); s->Ss.release();
    s->Xs = nx; s->Ys = ny; s->Ss = ns;
    s->caprows = cap;
  }
  launch_desc(s, d, c, s->Xs.p + s->nrows * din, false);
  std::vector<float> y(N);
  for (int i = 0; i < N; i++) y[i] = (float)labelsSlot[i];
  HIP_CHECK(hipMemcpyAsync(s->Ys.p + s->nrows, y.data(), sizeof(float) * N, hipMemcpyHostToDevice, d->stream));
  HIP_CHECK(hipMemcpyAsync(s->Ss.p + s->nrows, d->type.p, sizeof(int) * N, hipMemcpyDeviceToDevice, d->stream));
  std::vector<int> sp(N);
  int nover = 0;
  HIP_CHECK(hipMemcpyAsync(sp.data(), d->type.p, sizeof(int) * N, hipMemcpyDeviceToHost, d->stream));
  HIP_CHECK(hipMemcpyAsync(&nover, s->nearOver.p, sizeof(int), hipMemcpyDeviceToHost, d->stream));
  HIP_CHECK(hipStreamSynchronize(d->stream));
  if (nover) fatal(task, "more neighbours within the descriptor cutoff than the model holds (256)");
  s->hostY.insert(s->hostY.end(), y.begin(), y.end());
  s->hostS.insert(s->hostS.end(), sp.begin(), sp.end());
  s->nrows += N;
}

}  // namespace emdee

// ---- C ABI (include/emdee_gpu.h) ---------------------------------------------------

extern "C" {

EmDeeSurrogate* EmDeeSurrogate_create(int ntypes, double rcut, unsigned long long seed) {
  if (ntypes < 1 || ntypes > SG_MAXT) fatal("surrogate", "between 1 and 4 species are supported");
  EmDeeSurrogate* s = new EmDeeSurrogate();
  s->sp.nt = ntypes;
  s->sp.din = ntypes * SG_NB;
  s->sp.rc = (float)rcut;
  s->sp.mu0 = 0.5f;
  s->sp.dmu = (float)((rcut - 0.5) / (SG_NB - 1));
  s->sp.eta = 0.5f / (s->sp.dmu * s->sp.dmu);
  ParamLayout pl{ s->sp.din, s->hidden };
  s->P = pl.size();
  // Glorot-uniform weights from a splitmix64 stream, zero biases
  s->hostParams.assign((size_t)ntypes * s->P, 0.0f);
  uint64_t state = seed;
  auto next = [&]() {
    state += 0x9E3779B97F4A7C15ULL;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    z ^= z >> 31;
    return ((double)(z >> 11) + 0.5) / 9007199254740992.0;
  };
  for (int t = 0; t < ntypes; t++) {
    float* p = s->hostParams.data() + (size_t)t * s->P;
    double a1 = std::sqrt(6.0 / (pl.din + pl.H)), a2 = std::sqrt(6.0 / (2.0 * pl.H)), a3 = std::sqrt(6.0 / (pl.H + 1));
    for (int k = 0; k < pl.din * pl.H; k++) p[pl.W1() + k] = (float)(a1 * (2 * next() - 1));
    for (int k = 0; k < pl.H * pl.H; k++) p[pl.W2() + k] = (float)(a2 * (2 * next() - 1));
    for (int k = 0; k < pl.H; k++) p[pl.w3() + k] = (float)(a3 * (2 * next() - 1));
  }
  size_t np = (size_t)ntypes * s->P;
  s->params.alloc(np); s->grads.alloc(np); s->am.alloc(np); s->av.alloc(np);
  HIP_CHECK(hipMemcpy(s->params.p, s->hostParams.data(), sizeof(float) * np, hipMemcpyHostToDevice));
  HIP_CHECK(hipMemset(s->grads.p, 0, sizeof(float) * np));
  HIP_CHECK(hipMemset(s->am.p, 0, sizeof(float) * np));
  HIP_CHECK(hipMemset(s->av.p, 0, sizeof(float) * np));
  s->w1t.alloc((size_t)ntypes * pl.din * pl.H); s->w1.alloc((size_t)ntypes * pl.din * pl.H);
  s->w2t.alloc((size_t)ntypes * pl.H * pl.H); s->w2.alloc((size_t)ntypes * pl.H * pl.H);
  s->xmean.alloc(pl.din); s->xstd.alloc(pl.din);
  s->ymean.assign(ntypes, 0.0f); s->ystd.assign(ntypes, 1.0f);
  s->Fw.alloc(8);
  HIP_CHECK(hipStreamCreateWithFlags(&s->own, hipStreamNonBlocking));
  s->stream = s->own;
  repack(s);
  HIP_CHECK(hipStreamSynchronize(s->stream));
  return s;
}


// The exact long-range channel: the reciprocal Ewald sum over the k-vectors
// with |k| <= kcut (0 disables it).  Set before the first sample: the rows'
// labels are the residuals after it.
void EmDeeSurrogate_set_long_range(EmDeeSurrogate* s, double kcut) {
  if (kcut < 0.0) fatal("surrogate", "the long-range cutoff must not be negative");
  if (s->nrows > 0) fatal("surrogate", "the long-range channel must be set before the first sample");
  s->kcut = kcut;
  dev_lowk_release(s->lowk);
  s->lowk = nullptr;
  s->lowkOwner = nullptr;
}

int EmDeeSurrogate_long_range_vectors(EmDeeSurrogate* s) { return dev_lowk_vectors(s->lowk); }

int EmDeeSurrogate_rows(EmDeeSurrogate* s) { return (int)s->nrows; }

// One training pass (`epochs` epochs over the stored rows, `batchRows` rows
// per Adam step); returns the mean squared error of the normalised atom
// energies in the last epoch.  The first call fixes the normalisation.
double EmDeeSurrogate_train(EmDeeSurrogate* s, int epochs, int batchRows, double lr) {
  int din = s->sp.din, nt = s->sp.nt;
  long rows = s->nrows;
  if (rows == 0) fatal("surrogate training", "no training rows");
  s->stream = s->own;
  hipStream_t own = s->own;
  if (!s->normalised) {
    // feature means and deviations from the stored rows (host pass)
    std::vector<float> X((size_t)rows * din);
    HIP_CHECK(hipMemcpy(X.data(), s->Xs.p, sizeof(float) * X.size(), hipMemcpyDeviceToHost));
    std::vector<double> m(din, 0.0), v(din, 0.0);
    for (long r = 0; r < rows; r++) for (int f = 0; f < din; f++) m[f] += X[r * din + f];
    for (int f = 0; f < din; f++) m[f] /= rows;
    for (long r = 0; r < rows; r++) for (int f = 0; f < din; f++) { double dd = X[r * din + f] - m[f]; v[f] += dd * dd; }
    std::vector<float> mf(din), sf(din);
    for (int f = 0; f < din; f++) { mf[f] = (float)m[f]; sf[f] = (float)std::max(std::sqrt(v[f] / rows), 1e-3); }
    HIP_CHECK(hipMemcpy(s->xmean.p, mf.data(), sizeof(float) * din, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(s->xstd.p, sf.data(), sizeof(float) * din, hipMemcpyHostToDevice));
    for (int t = 0; t < nt; t++) {
      double sm = 0, sq = 0;
      long c = 0;
      for (long r = 0; r < rows; r++) if (s->hostS[r] == t) { sm += s->hostY[r]; c++; }
      double mean = c ? sm / c : 0.0;
      for (long r = 0; r < rows; r++) if (s->hostS[r] == t) { double dd = s->hostY[r] - mean; sq += dd * dd; }
      s->ymean[t] = (float)mean;
      s->ystd[t] = (float)std::max(c ? std::sqrt(sq / c) : 1.0, 1e-12);
    }
    s->normalised = true;
  }
  // normalised labels, per row: (y - ymean_s) / ystd_s
  std::vector<float> yn(rows);
  for (long r = 0; r < rows; r++) yn[r] = (s->hostY[r] - s->ymean[s->hostS[r]]) / s->ystd[s->hostS[r]];
  DBuf<float> Yn;
  Yn.alloc(rows);
  HIP_CHECK(hipMemcpy(Yn.p, yn.data(), sizeof(float) * rows, hipMemcpyHostToDevice));
  long batch = std::min<long>(std::max(batchRows, 64), rows);
  ensure_work(s, batch);
  double lastLoss = 0;
  for (int ep = 0; ep < epochs; ep++) {
    double lsum = 0;
    long lcount = 0;
    for (long r0 = 0; r0 < rows; r0 += batch) {
      long nr = std::min(batch, rows - r0);
      long pad = ((nr + SG_ROWS - 1) / SG_ROWS) * SG_ROWS;
      hipLaunchKernelGGL(k_sg_normalise, dim3(div_up(pad * din, BLOCK)), dim3(BLOCK), 0, own, s->X.p,
                         s->Xs.p + r0 * din, s->xmean.p, s->xstd.p, nr, din, pad);
      for (int t = 0; t < nt; t++) {
        long cnt = 0;
        for (long r = r0; r < r0 + nr; r++) cnt += s->hostS[r] == t;
        if (cnt == 0) continue;
        lsum += mlp_pass(s, t, nr, s->Ss.p + r0, Yn.p + r0, 1.0f / (float)cnt, 1, nullptr);
        lcount += cnt;
        ParamLayout pl{ din, s->hidden };
        s->adamStep++;
        float c1 = (float)(1.0 - std::pow(0.9, (double)s->adamStep)), c2 = (float)(1.0 - std::pow(0.999, (double)s->adamStep));
        hipLaunchKernelGGL(k_sg_adam, dim3(div_up(s->P, BLOCK)), dim3(BLOCK), 0, own, s->params.p + (size_t)t * s->P,
                           s->am.p + (size_t)t * s->P, s->av.p + (size_t)t * s->P, s->grads.p + (size_t)t * s->P, s->P,
                           (float)lr, 0.9f, 0.999f, 1e-8f, c1, c2);
        (void)pl;
      }
      repack(s);
    }
    lastLoss = lcount ? lsum / lcount : 0.0;
  }
  HIP_CHECK(hipStreamSynchronize(own));
  Yn.release();
  return lastLoss;
}

// Parameters: nt blocks of EmDeeSurrogate_param_count(s) floats, each laid
// out as W1[din][H], b1[H], W2[H][H], b2[H], w3[H], b3; normalisation as
// xmean[din], xstd[din], ymean[nt], ystd[nt].
int EmDeeSurrogate_param_count(EmDeeSurrogate* s) { return s->P; }
void EmDeeSurrogate_get_params(EmDeeSurrogate* s, float* params, float* norm) {
  HIP_CHECK(hipMemcpy(params, s->params.p, sizeof(float) * s->P * s->sp.nt, hipMemcpyDeviceToHost));
  if (norm) {
    int din = s->sp.din;
    HIP_CHECK(hipMemcpy(norm, s->xmean.p, sizeof(float) * din, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(norm + din, s->xstd.p, sizeof(float) * din, hipMemcpyDeviceToHost));
    for (int t = 0; t < s->sp.nt; t++) { norm[2 * din + t] = s->ymean[t]; norm[2 * din + s->sp.nt + t] = s->ystd[t]; }
  }
}
void EmDeeSurrogate_set_params(EmDeeSurrogate* s, const float* params, const float* norm) {
  s->stream = s->own;
  HIP_CHECK(hipMemcpy(s->params.p, params, sizeof(float) * s->P * s->sp.nt, hipMemcpyHostToDevice));
  if (norm) {
    int din = s->sp.din;
    HIP_CHECK(hipMemcpy(s->xmean.p, norm, sizeof(float) * din, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(s->xstd.p, norm + din, sizeof(float) * din, hipMemcpyHostToDevice));
    for (int t = 0; t < s->sp.nt; t++) { s->ymean[t] = norm[2 * din + t]; s->ystd[t] = norm[2 * din + s->sp.nt + t]; }
    s->normalised = true;
  }
  repack(s);
  HIP_CHECK(hipStreamSynchronize(s->stream));
}

void EmDeeSurrogate_destroy(EmDeeSurrogate* s) {
  if (!s) return;
  for (DBuf<float>* b : { &s->params, &s->grads, &s->am, &s->av, &s->xmean, &s->xstd, &s->Xs, &s->Ys, &s->X, &s->Z1,
                          &s->H1, &s->Z2, &s->H2, &s->G2, &s->G1, &s->dX, &s->e, &s->dXsum, &s->part })
    b->release();
  s->w1t.release(); s->w1.release(); s->w2t.release(); s->w2.release(); s->Ss.release();
  s->cJ.release(); s->cN.release(); s->cOver.release(); s->rowSlot.release(); s->grpSpecies.release(); s->Rf.release();
  s->nearE.release(); s->nearG.release(); s->nearN.release(); s->nearOver.release();
  s->Xi.release(); s->dEdD.release(); s->eat.release(); s->epart.release(); s->wpart.release();
  s->lowkAtom.release();
  dev_lowk_release(s->lowk);
  s->Fw.release(); s->red.release();
  (void)hipStreamDestroy(s->own);
  delete s;
}

}  // extern "C"
