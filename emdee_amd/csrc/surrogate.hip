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
#include "surrogate.h"

namespace emdee {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
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
  // work buffers (rows padded to SG_ROWS)
  long workRows = 0;
  DBuf<float> X, Z1, H1, Z2, H2, G2, G1, dX, e, dXsum, part;
  DBuf<double> Fw, red;          // results; per-block partials
  DBuf<float4> lE;               // compact neighbour list (k_sg_list)
  DBuf<int> lJ, lN, lOver;
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

__device__ inline void pair_sep(const double* __restrict__ R, int Npad, int i, int j, double L, double invL,
                                float& dx, float& dy, float& dz) {
  double ex = R[i] - R[j], ey = R[Npad + i] - R[Npad + j], ez = R[2 * Npad + i] - R[2 * Npad + j];
  ex -= L * rint(ex * invL);
  ey -= L * rint(ey * invL);
  ez -= L * rint(ez * invL);
  dx = (float)ex; dy = (float)ey; dz = (float)ez;
}

// Compact neighbour list within rc, built from the engine's Verlet list at
// every use: SG_SUBS consecutive lanes per atom scan interleaved parts of its
// list and each appends its hits to its own segment; an entry keeps the
// separation vector and distance (FP32) and the neighbour's slot with its
// species in the low two bits.  Segments are interleaved: entry c of segment
// s of atom i sits at (i SG_SEG + c) SG_SUBS + s, so the lanes of one atom
// write and read neighbouring addresses.
constexpr int SG_SUBS = 16, SG_SEG = 40;
__device__ inline size_t sg_at(int i, int c, int s) { return ((size_t)i * SG_SEG + c) * SG_SUBS + s; }
__global__ void __launch_bounds__(BLOCK) k_sg_list(float4* __restrict__ E, int* __restrict__ J, int* __restrict__ cnt,
                                                   int* __restrict__ overflow, const double* __restrict__ R,
                                                   const int* __restrict__ nbr, const int* __restrict__ nnb,
                                                   const int* __restrict__ type, float rc, double L, int N,
                                                   int Npad) {
  int gt = blockIdx.x * blockDim.x + threadIdx.x;
  int i = gt / SG_SUBS, sub = gt % SG_SUBS;
  if (i >= N) return;
  const float rc2 = rc * rc;
  const double invL = 1.0 / L;
  int n = nnb[i], c = 0;
  for (int k = sub; k < n; k += SG_SUBS) {
    int j = nbr[(size_t)k * Npad + i];
    float dx, dy, dz;
    pair_sep(R, Npad, i, j, L, invL, dx, dy, dz);
    float r2 = dx * dx + dy * dy + dz * dz;
    if (r2 < rc2) {
      if (c < SG_SEG) { E[sg_at(i, c, sub)] = make_float4(dx, dy, dz, sqrtf(r2)); J[sg_at(i, c, sub)] = (j << 2) | type[j]; }
      c++;
    }
  }
  cnt[i * SG_SUBS + sub] = min(c, SG_SEG);
  if (c > SG_SEG) atomicMax(overflow, c);
}

// Descriptors from the compact list: 32 lanes per atom, lane b owns Gaussian b
// of every neighbour species; entries are read once per atom (broadcast).
template <int NT>
__global__ void __launch_bounds__(BLOCK) k_sg_desc2(float* __restrict__ D, const float4* __restrict__ E,
                                                    const int* __restrict__ J, const int* __restrict__ cnt,
                                                    const int* __restrict__ type, const float* __restrict__ mean,
                                                    const float* __restrict__ sd, SurrogateParams sp, int N) {
  int gt = blockIdx.x * blockDim.x + threadIdx.x;
  int i = gt / SG_NB, b = gt % SG_NB;
  if (i >= N) return;
  float acc[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) acc[t] = 0.0f;
  const float mu = sp.mu0 + b * sp.dmu, kpi = PI_F / sp.rc;
  for (int s = 0; s < SG_SUBS; s++) {
    int n = cnt[i * SG_SUBS + s];
    for (int c = 0; c < n; c++) {
      float r = E[sg_at(i, c, s)].w;
      int t = J[sg_at(i, c, s)] & 3;
      float u = r - mu, g = __expf(-sp.eta * u * u) * 0.5f * (__cosf(kpi * r) + 1.0f);
#pragma unroll
      for (int q = 0; q < NT; q++) acc[q] += (q == t) ? g : 0.0f;
    }
  }
#pragma unroll
  for (int t = 0; t < NT; t++) {
    int f = t * SG_NB + b;
    D[(size_t)i * (NT * SG_NB) + f] = mean ? (acc[t] - mean[f]) / sd[f] : acc[t];   // normalised when mean given
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

// Forces from dE/dD over the compact list: 16 lanes per atom share its
// entries; each entry sums its 32 Gaussian derivatives; the lanes' forces are
// combined by a fixed butterfly.
template <int NT>
__global__ void __launch_bounds__(BLOCK) k_sg_force2(double* __restrict__ F, double* __restrict__ red,
                                                     const float* __restrict__ dXn, const float* __restrict__ xstd,
                                                     const float4* __restrict__ E, const int* __restrict__ J,
                                                     const int* __restrict__ cnt, const int* __restrict__ type,
                                                     SurrogateParams sp, int N, int Npad) {
  constexpr int TPA = 16, din = NT * SG_NB;
  int gt = blockIdx.x * blockDim.x + threadIdx.x;
  int i = gt / TPA, sub = gt % TPA;
  double fx = 0, fy = 0, fz = 0, w[1] = { 0.0 };
  static_assert(TPA == SG_SUBS, "one lane per list segment");
  if (i < N) {
    const float kpi = PI_F / sp.rc;
    int ti = type[i];
    {
      int n = cnt[i * SG_SUBS + sub];
      for (int c = 0; c < n; c++) {
        float4 e = E[sg_at(i, c, sub)];
        int jt = J[sg_at(i, c, sub)], j = jt >> 2, tj = jt & 3;
        float r = e.w, fc = 0.5f * (__cosf(kpi * r) + 1.0f), dfc = -0.5f * kpi * __sinf(kpi * r);
        const float* gi = dXn + (size_t)i * din + tj * SG_NB;
        const float* gj = dXn + (size_t)j * din + ti * SG_NB;
        const float* si = xstd + tj * SG_NB;
        const float* sj = xstd + ti * SG_NB;
        float dEdr = 0.0f;
#pragma unroll 8
        for (int b = 0; b < SG_NB; b++) {
          float u = r - (sp.mu0 + b * sp.dmu), g = __expf(-sp.eta * u * u);
          dEdr += (gi[b] / si[b] + gj[b] / sj[b]) * (g * (dfc - 2.0f * sp.eta * u * fc));
        }
        float k = -dEdr / r;
        fx += (double)(k * e.x);
        fy += (double)(k * e.y);
        fz += (double)(k * e.z);
        w[0] += -0.5 * (double)(dEdr * r);
      }
    }
  }
#pragma unroll
  for (int off = TPA / 2; off > 0; off >>= 1) {
    fx += __shfl_xor(fx, off, 64);
    fy += __shfl_xor(fy, off, 64);
    fz += __shfl_xor(fz, off, 64);
  }
  if (i < N && sub == 0) {
    F[i] += fx;
    F[Npad + i] += fy;
    F[2 * Npad + i] += fz;
  }
  block_sum<1>(w);
  if (threadIdx.x == 0) red[blockIdx.x] = w[0];
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

static void build_list(EmDeeSurrogate* s, Device* d, double L, hipStream_t st) {
  int N = d->N;
  s->lE.alloc((size_t)N * SG_SUBS * SG_SEG);
  s->lJ.alloc((size_t)N * SG_SUBS * SG_SEG);
  s->lN.alloc((size_t)N * SG_SUBS);
  s->lOver.alloc(1);
  HIP_CHECK(hipMemsetAsync(s->lOver.p, 0, sizeof(int), st));
  hipLaunchKernelGGL(k_sg_list, dim3(div_up((long)d->Npad * SG_SUBS, BLOCK)), dim3(BLOCK), 0, st, s->lE.p, s->lJ.p, s->lN.p,
                     s->lOver.p, d->R.p, d->nbr.p, d->nnb.p, d->type.p, s->sp.rc, L, N, d->Npad);
  launch_check("surrogate list");
  int over = 0;
  HIP_CHECK(hipMemcpyAsync(&over, s->lOver.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipStreamSynchronize(st));
  if (over) fatal("surrogate", "more neighbours within the descriptor cutoff than the compact list holds");
}

static void launch_desc(EmDeeSurrogate* s, Device* d, float* D, double L, hipStream_t st, bool normalised) {
  build_list(s, d, L, st);
  int g = div_up((long)d->N * SG_NB, BLOCK);
  const float* m = normalised ? s->xmean.p : nullptr;
  const float* sd = normalised ? s->xstd.p : nullptr;
#define SG_DESC(NT) hipLaunchKernelGGL(k_sg_desc2<NT>, dim3(g), dim3(BLOCK), 0, st, D, s->lE.p, s->lJ.p, s->lN.p, d->type.p, m, sd, s->sp, d->N)
  switch (s->sp.nt) { case 1: SG_DESC(1); break; case 2: SG_DESC(2); break; case 3: SG_DESC(3); break; default: SG_DESC(4); }
#undef SG_DESC
  launch_check("surrogate descriptors");
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

// Inference on a system's device state: adds the model forces to F (slot
// order) and returns the model energy and pair virial; eatom (slot order, may
// be null) receives the atom energies.
__global__ void k_sg_add_atom(double* __restrict__ dst, const float* __restrict__ src, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] += (double)src[i];
}

double dev_surrogate_apply(Device* d, EmDeeSurrogate* s, int layer, double L, double* W) {
  if (!s->normalised) fatal("surrogate", "the model has not been trained");
  if (s->sp.nt != d->ntypes) fatal("surrogate", "the model and the system have different numbers of types");
  int N = d->N;
  double* F = d->F.p + 3 * (size_t)d->Npad * layer;
  if (N == 0) { *W = 0; return 0; }
  s->stream = d->stream;
  ensure_work(s, N);
  long pad = s->workRows;
  int din = s->sp.din;
  // normalised descriptors straight into X (rows past N stay zero)
  launch_desc(s, d, s->X.p, L, s->stream, true);
  double E = 0;
  float* eatom = s->e.p;
  for (int t = 0; t < s->sp.nt; t++) E += mlp_pass(s, t, N, d->type.p, nullptr, 0.0f, 0, eatom);
  int g = div_up((long)N * 16, BLOCK);
#define SG_FORCE(NT)                                                                                                \
  hipLaunchKernelGGL(k_sg_force2<NT>, dim3(g), dim3(BLOCK), 0, s->stream, F, s->red.p, s->dXsum.p, s->xstd.p, s->lE.p, \
                     s->lJ.p, s->lN.p, d->type.p, s->sp, N, d->Npad)
  switch (s->sp.nt) { case 1: SG_FORCE(1); break; case 2: SG_FORCE(2); break; case 3: SG_FORCE(3); break; default: SG_FORCE(4); }
#undef SG_FORCE
  launch_check("surrogate forces");
  hipLaunchKernelGGL(k_reduce_partials<1>, dim3(1), dim3(BLOCK), 0, s->stream, s->Fw.p + 1, s->red.p, g);
  HIP_CHECK(hipMemcpyAsync(W, s->Fw.p + 1, sizeof(double), hipMemcpyDeviceToHost, s->stream));
  if (d->wantAtomE)
    hipLaunchKernelGGL(k_sg_add_atom, dim3(div_up(N, BLOCK)), dim3(BLOCK), 0, s->stream, d->eatom.p, eatom, N);
  HIP_CHECK(hipStreamSynchronize(s->stream));
  return E;
}

// Stores the raw descriptors of a system's current configuration as training
// rows, with per-atom labels given in slot order.
void dev_surrogate_add_sample(Device* d, EmDeeSurrogate* s, double L, const double* labels) {
  int N = d->N, din = s->sp.din;
  s->stream = d->stream;
  std::vector<int> iperm(N);
  HIP_CHECK(hipMemcpyAsync(iperm.data(), d->iperm.p, sizeof(int) * N, hipMemcpyDeviceToHost, d->stream));
  HIP_CHECK(hipStreamSynchronize(d->stream));
  std::vector<double> labelsSlot(N);
  for (int a = 0; a < N; a++) labelsSlot[iperm[a]] = labels[a];
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
  launch_desc(s, d, s->Xs.p + s->nrows * din, L, d->stream, false);
  std::vector<float> y(N);
  for (int i = 0; i < N; i++) y[i] = (float)labelsSlot[i];
  HIP_CHECK(hipMemcpyAsync(s->Ys.p + s->nrows, y.data(), sizeof(float) * N, hipMemcpyHostToDevice, d->stream));
  HIP_CHECK(hipMemcpyAsync(s->Ss.p + s->nrows, d->type.p, sizeof(int) * N, hipMemcpyDeviceToDevice, d->stream));
  std::vector<int> sp(N);
  HIP_CHECK(hipMemcpyAsync(sp.data(), d->type.p, sizeof(int) * N, hipMemcpyDeviceToHost, d->stream));
  HIP_CHECK(hipStreamSynchronize(d->stream));
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
  s->lE.release(); s->lJ.release(); s->lN.release(); s->lOver.release();
  s->Fw.release(); s->red.release();
  (void)hipStreamDestroy(s->own);
  delete s;
}

}  // extern "C"
