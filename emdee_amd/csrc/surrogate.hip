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

__device__ inline void pair_sep(const double* __restrict__ R, int Npad, int i, int j, double L, float& dx, float& dy,
                                float& dz) {
  double ex = R[i] - R[j], ey = R[Npad + i] - R[Npad + j], ez = R[2 * Npad + i] - R[2 * Npad + j];
  ex -= L * rint(ex / L);
  ey -= L * rint(ey / L);
  ez -= L * rint(ez / L);
  dx = (float)ex; dy = (float)ey; dz = (float)ez;
}

// Raw descriptors of every slot: D[slot][t*NB + k].
template <int NT>
__global__ void __launch_bounds__(BLOCK) k_sg_desc(float* __restrict__ D, const double* __restrict__ R,
                                                   const int* __restrict__ type, const int* __restrict__ nbr,
                                                   const int* __restrict__ nnb, SurrogateParams sp, double L, int N,
                                                   int Npad) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  float acc[NT * SG_NB];
#pragma unroll
  for (int f = 0; f < NT * SG_NB; f++) acc[f] = 0.0f;
  const float rc2 = sp.rc * sp.rc;
  int n = nnb[i];
  for (int k = 0; k < n; k++) {
    int j = nbr[(size_t)k * Npad + i];
    float dx, dy, dz;
    pair_sep(R, Npad, i, j, L, dx, dy, dz);
    float r2 = dx * dx + dy * dy + dz * dz;
    if (r2 >= rc2) continue;
    float r = sqrtf(r2), fc = 0.5f * (__cosf(PI_F * r / sp.rc) + 1.0f);
    int t = type[j];
#pragma unroll
    for (int q = 0; q < NT; q++) {
      if (q != t) continue;
#pragma unroll
      for (int b = 0; b < SG_NB; b++) {
        float u = r - (sp.mu0 + b * sp.dmu);
        acc[q * SG_NB + b] += __expf(-sp.eta * u * u) * fc;
      }
    }
  }
#pragma unroll
  for (int f = 0; f < NT * SG_NB; f++) D[(size_t)i * (NT * SG_NB) + f] = acc[f];
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
template <int K, int NC, bool ACT>
__global__ void __launch_bounds__(256) k_sg_gemm(const float* __restrict__ A, const unsigned short* __restrict__ Bt,
                                                 const float* __restrict__ bias, float* __restrict__ Z,
                                                 float* __restrict__ Hout) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const long r0 = (long)blockIdx.x * SG_ROWS + 16 * w;
  const int ai = lane & 15, ak = 4 * (lane >> 4);
  f32x4 acc[NC / 16];
#pragma unroll
  for (int t = 0; t < NC / 16; t++) acc[t] = f32x4{ 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
  for (int k0 = 0; k0 < K; k0 += 16) {
    float4 av = *reinterpret_cast<const float4*>(A + (r0 + ai) * K + k0 + ak);
    bf16x4 a = { (short)f2bf(av.x), (short)f2bf(av.y), (short)f2bf(av.z), (short)f2bf(av.w) };
#pragma unroll
    for (int t = 0; t < NC / 16; t++) {
      bf16x4 b = *reinterpret_cast<const bf16x4*>(Bt + (size_t)(16 * t + ai) * K + k0 + ak);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int t = 0; t < NC / 16; t++) {
    int j = 16 * t + ai;
    float bj = bias ? bias[j] : 0.0f;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      long i = r0 + ak + r;
      float z = acc[t][r] + bj;
      Z[i * NC + j] = z;
      if (ACT) Hout[i * NC + j] = silu(z);
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

// Output layer of species s and the seed of the backward pass.
//   mode 0 (inference): eatom[row] = ystd out + ymean for rows of species s,
//       seed = ystd (dE/d out).
//   mode 1 (training): loss partials of (out - y)^2, seed = 2 (out - y) / count.
// G2 = seed * w3 * SiLU'(Z2) (zero for rows of other species).
__global__ void k_sg_output(float* __restrict__ G2, float* __restrict__ eatom, float* __restrict__ seedOut,
                            double* __restrict__ red, const float* __restrict__ H2, const float* __restrict__ Z2,
                            const float* __restrict__ p, const int* __restrict__ species, const float* __restrict__ y,
                            int s, float ymean, float ystd, float invCount, long rows, int mode, ParamLayout pl) {
  long row = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double v[1] = { 0.0 };
  if (row < rows) {
    const float* w3 = p + pl.w3();
    float out = p[pl.b3()];
    for (int n = 0; n < SG_H; n++) out += H2[row * SG_H + n] * w3[n];
    bool mine = species[row] == s;
    float seed = 0.0f;
    if (mine) {
      if (mode == 0) {
        eatom[row] = ystd * out + ymean;
        seed = ystd;
        v[0] = (double)eatom[row];
      } else {
        float d = out - y[row];
        seed = 2.0f * d * invCount;
        v[0] = (double)d * (double)d;
      }
    }
    seedOut[row] = seed;
    for (int n = 0; n < SG_H; n++) G2[row * SG_H + n] = seed * w3[n] * dsilu(Z2[row * SG_H + n]);
  }
  block_sum<1>(v);
  if (threadIdx.x == 0) red[blockIdx.x] = v[0];
}

__global__ void k_sg_dsilu(float* __restrict__ G, const float* __restrict__ Z, long n) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < n) G[idx] *= dsilu(Z[idx]);
}
__global__ void k_sg_accumulate(float* __restrict__ dst, const float* __restrict__ src, long n, int first) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < n) dst[idx] = first ? src[idx] : dst[idx] + src[idx];
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

// Forces from dE/dD (normalised-input gradient dXn / xstd) and the model's
// pair virial; F (slot order, SoA) is incremented.
template <int NT>
__global__ void __launch_bounds__(BLOCK) k_sg_force(double* __restrict__ F, double* __restrict__ red,
                                                    const float* __restrict__ dXn, const float* __restrict__ xstd,
                                                    const double* __restrict__ R, const int* __restrict__ type,
                                                    const int* __restrict__ nbr, const int* __restrict__ nnb,
                                                    SurrogateParams sp, double L, int N, int Npad) {
  const int din = NT * SG_NB;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double w[1] = { 0.0 };
  if (i < N) {
    const float rc2 = sp.rc * sp.rc, kpi = PI_F / sp.rc;
    int ti = type[i];
    double fx = 0, fy = 0, fz = 0;
    int n = nnb[i];
    for (int k = 0; k < n; k++) {
      int j = nbr[(size_t)k * Npad + i];
      float dx, dy, dz;
      pair_sep(R, Npad, i, j, L, dx, dy, dz);
      float r2 = dx * dx + dy * dy + dz * dz;
      if (r2 >= rc2) continue;
      float r = sqrtf(r2);
      float fc = 0.5f * (__cosf(kpi * r) + 1.0f), dfc = -0.5f * kpi * __sinf(kpi * r);
      int tj = type[j];
      const float* gi = dXn + (size_t)i * din + tj * SG_NB;
      const float* gj = dXn + (size_t)j * din + ti * SG_NB;
      const float* si = xstd + tj * SG_NB;
      const float* sj = xstd + ti * SG_NB;
      float dEdr = 0.0f;
#pragma unroll 8
      for (int b = 0; b < SG_NB; b++) {
        float u = r - (sp.mu0 + b * sp.dmu), g = __expf(-sp.eta * u * u);
        float dG = g * (dfc - 2.0f * sp.eta * u * fc);
        dEdr += (gi[b] / si[b] + gj[b] / sj[b]) * dG;
      }
      float c = -dEdr / r;
      fx += (double)(c * dx);
      fy += (double)(c * dy);
      fz += (double)(c * dz);
      w[0] += -0.5 * (double)(dEdr * r);
    }
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
  s->red.alloc(pad / 64 + 64);   // >= one partial per 256-row block
  s->part.alloc((size_t)(pad / 1024 + 2) * H * H);
  HIP_CHECK(hipMemsetAsync(s->X.p, 0, sizeof(float) * pad * din, s->stream));
  for (DBuf<float>* b : { &s->Z1, &s->H1, &s->Z2, &s->H2, &s->G2, &s->G1 })
    HIP_CHECK(hipMemsetAsync(b->p, 0, sizeof(float) * pad * H, s->stream));
  s->workRows = pad;
}

static void launch_desc(EmDeeSurrogate* s, Device* d, float* D, double L, hipStream_t st) {
  int g = div_up(d->N, BLOCK);
  switch (s->sp.nt) {
    case 1: hipLaunchKernelGGL(k_sg_desc<1>, dim3(g), dim3(BLOCK), 0, st, D, d->R.p, d->type.p, d->nbr.p, d->nnb.p, s->sp, L, d->N, d->Npad); break;
    case 2: hipLaunchKernelGGL(k_sg_desc<2>, dim3(g), dim3(BLOCK), 0, st, D, d->R.p, d->type.p, d->nbr.p, d->nnb.p, s->sp, L, d->N, d->Npad); break;
    case 3: hipLaunchKernelGGL(k_sg_desc<3>, dim3(g), dim3(BLOCK), 0, st, D, d->R.p, d->type.p, d->nbr.p, d->nnb.p, s->sp, L, d->N, d->Npad); break;
    default: hipLaunchKernelGGL(k_sg_desc<4>, dim3(g), dim3(BLOCK), 0, st, D, d->R.p, d->type.p, d->nbr.p, d->nnb.p, s->sp, L, d->N, d->Npad); break;
  }
  launch_check("surrogate descriptors");
}

static void gemm(EmDeeSurrogate* s, int K, int NC, bool act, const float* A, const unsigned short* Bt, const float* bias,
                 float* Z, float* H, long padRows) {
  dim3 grid((unsigned)(padRows / SG_ROWS));
#define SG_GEMM(KK, NN, AA) hipLaunchKernelGGL((k_sg_gemm<KK, NN, AA>), grid, dim3(256), 0, s->stream, A, Bt, bias, Z, H)
  if (K == 64 && NC == 128) { if (act) SG_GEMM(64, 128, true); else SG_GEMM(64, 128, false); }
  else if (K == 128 && NC == 128) { if (act) SG_GEMM(128, 128, true); else SG_GEMM(128, 128, false); }
  else if (K == 128 && NC == 64) SG_GEMM(128, 64, false);
  else if (K == 32 && NC == 128) { if (act) SG_GEMM(32, 128, true); else SG_GEMM(32, 128, false); }
  else if (K == 128 && NC == 32) SG_GEMM(128, 32, false);
  else fatal("surrogate", "unsupported layer shape");
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
  gemm(s, din, H, true, s->X.p, w1t, p + pl.b1(), s->Z1.p, s->H1.p, pad);
  gemm(s, H, H, true, s->H1.p, w2t, p + pl.b2(), s->Z2.p, s->H2.p, pad);
  int gr = div_up(rows, BLOCK);
  float* seed = s->e.p + pad;   // second half of e holds the per-row seeds
  hipLaunchKernelGGL(k_sg_output, dim3(gr), dim3(BLOCK), 0, s->stream, s->G2.p, eatom, seed, s->red.p, s->H2.p, s->Z2.p, p,
                     species, y, spc, s->ymean[spc], s->ystd[spc], invCount, rows, mode, pl);
  launch_check("surrogate output");
  // G1 = (G2 W2^T) * SiLU'(Z1)
  gemm(s, H, H, false, s->G2.p, w2, nullptr, s->G1.p, nullptr, pad);
  hipLaunchKernelGGL(k_sg_dsilu, dim3(div_up(pad * H, BLOCK)), dim3(BLOCK), 0, s->stream, s->G1.p, s->Z1.p, pad * H);
  if (mode == 0) {
    gemm(s, H, din, false, s->G1.p, w1, nullptr, s->dX.p, nullptr, pad);
  } else {
    // weight gradients of this species into grads
    float* g = s->grads.p + (size_t)spc * s->P;
    int rpb = 1024, nb = (int)((rows + rpb - 1) / rpb);
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
  // descriptors into dX (scratch), then normalised into X
  launch_desc(s, d, s->dX.p, L, s->stream);
  hipLaunchKernelGGL(k_sg_normalise, dim3(div_up(pad * din, BLOCK)), dim3(BLOCK), 0, s->stream, s->X.p, s->dX.p,
                     s->xmean.p, s->xstd.p, (long)N, din, pad);
  double E = 0;
  float* eatom = s->e.p;
  for (int t = 0; t < s->sp.nt; t++) {
    E += mlp_pass(s, t, N, d->type.p, nullptr, 0.0f, 0, eatom);
    hipLaunchKernelGGL(k_sg_accumulate, dim3(div_up(pad * din, BLOCK)), dim3(BLOCK), 0, s->stream, s->dXsum.p, s->dX.p,
                       pad * din, t == 0 ? 1 : 0);
  }
  int g = div_up(N, BLOCK);
#define SG_FORCE(NT)                                                                                              \
  hipLaunchKernelGGL(k_sg_force<NT>, dim3(g), dim3(BLOCK), 0, s->stream, F, s->red.p, s->dXsum.p, s->xstd.p, d->R.p, \
                     d->type.p, d->nbr.p, d->nnb.p, s->sp, L, N, d->Npad)
  switch (s->sp.nt) { case 1: SG_FORCE(1); break; case 2: SG_FORCE(2); break; case 3: SG_FORCE(3); break; default: SG_FORCE(4); }
#undef SG_FORCE
  launch_check("surrogate forces");
  hipLaunchKernelGGL(k_reduce_partials<1>, dim3(1), dim3(BLOCK), 0, s->stream, s->Fw.p + 1, s->red.p, g);
  HIP_CHECK(hipMemcpyAsync(W, s->Fw.p + 1, sizeof(double), hipMemcpyDeviceToHost, s->stream));
  if (d->wantAtomE)
    hipLaunchKernelGGL(k_sg_add_atom, dim3(g), dim3(BLOCK), 0, s->stream, d->eatom.p, eatom, N);
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
  launch_desc(s, d, s->Xs.p + s->nrows * din, L, d->stream);
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
  s->Fw.release(); s->red.release();
  (void)hipStreamDestroy(s->own);
  delete s;
}

}  // extern "C"
