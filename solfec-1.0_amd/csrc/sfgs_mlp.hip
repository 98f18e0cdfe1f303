// sfgs_mlp.hip -- learned warm start for the Gauss-Seidel solve.
//
// A multilayer perceptron 32 -> 128 -> 128 -> 128 -> 32 (18 input features and
// 3 outputs used, the rest zero padding), GELU on the hidden layers, bf16
// weights and activations, FP32 accumulation.  It proposes each contact's
// reaction as R0 = R_prev + s * f(features); the plain Gauss-Seidel solve then
// runs from R0 as usual, so the network only changes where the solve starts.
//
// Every matrix product (inference, the backward pass, the weight gradients)
// goes through one tiled GEMM on the gfx950 matrix cores,
// v_mfma_f32_16x16x32_bf16: a 256-thread workgroup computes a 64x64 tile of C
// in four 32x32 wave tiles, staging 64x32 slices of A and B in LDS.  Lane l of
// a wave holds, for one 16x16x32 product, A[l%16][8(l/16)..+7] and
// B[8(l/16)..+7][l%16] and receives D[4(l/16)..+3][l%16].  Only the row and
// column placement matters for the result: A and B give their k values to the
// same lanes in the same order, so any assignment of the 32 k values to
// (lane group, element) is a permutation of the contraction and sums the same.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <cmath>
#include <cstring>
#include <vector>
#include <algorithm>
#include "sfgs_ctx.hpp"

namespace sfgs_mlp {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

static const int BM = 64, BN = 64, BK = 32, LPAD = 8;

enum Epi { EPI_F32 = 0, EPI_BIAS_GELU = 1, EPI_BIAS_F32 = 2, EPI_GELU_BWD = 3, EPI_BF16 = 4 };

// GELU with erf from Abramowitz & Stegun, Handbook of Mathematical Functions,
// eq. 7.1.26: erf(u) = 1 - t (a1 + t (a2 + t (a3 + t (a4 + t a5)))) exp(-u^2),
// t = 1 / (1 + p u), u >= 0, |error| <= 1.5e-7 (odd extension for u < 0).  The
// activation's absolute error, 0.75e-7 |x| at most, is far below the bf16
// resolution of the activations, and every kernel here uses this one form, so
// the fused inference and the layer-by-layer forward pass agree bit for bit.
// The exponential also gives the derivative's Gaussian: exp(-u^2) = exp(-x^2/2).
__device__ __forceinline__ void erf_as(float x, float &erf_u, float &gauss)
{
  const float u = fabsf(x)*0.70710678118654752f;
  const float t = __builtin_amdgcn_rcpf(1.0f + 0.3275911f*u);   // hardware reciprocal (1 ulp)
  const float poly = t*(0.254829592f + t*(-0.284496736f + t*(1.421413741f + t*(-1.453152027f + t*1.061405429f))));
  gauss = __expf(-u*u);
  const float e = 1.0f - poly*gauss;
  erf_u = x < 0.0f ? -e : e;
}
__device__ __forceinline__ float gelu(float x)
{
  float e, g;
  erf_as(x, e, g);
  return 0.5f*x*(1.0f + e);
}
__device__ __forceinline__ float dgelu(float x)
{
  float e, g;
  erf_as(x, e, g);
  return 0.5f*(1.0f + e) + x*0.39894228040143268f*g;
}

struct GemmArgs {
  int M, N, K;           // C is M x N, contraction over K
  const bf16 *A; int lda; int transA;   // transA = 0: A[m*lda + k]; 1: A[k*lda + m]
  const bf16 *B; int ldb; int transB;   // transB = 0: B[k*ldb + n]; 1: B[n*ldb + k]
  float *C; int ldc;                    // FP32 output (EPI_F32, EPI_BIAS_F32, EPI_BIAS_GELU keeps Z here)
  bf16 *H; int ldh;                     // bf16 output (EPI_BIAS_GELU: GELU(Z); EPI_GELU_BWD: dZ; EPI_BF16)
  const float *bias;                    // length N (EPI_BIAS_*)
  const float *Z; int ldz;              // pre-activations for EPI_GELU_BWD
  int epi;
  int kslice;            // split-K: workgroup z takes K range [z kslice, (z+1) kslice) and writes
                         // its partial C at C + z M ldc (EPI_F32 only); 0 = whole K
};

// eight consecutive bf16 of one operand row (16-byte aligned when ld and the
// base are multiples of 8 elements, which every operand here is)
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ u32x4 ld8(const bf16 *p) { return *reinterpret_cast<const u32x4*>(p); }
__device__ __forceinline__ bf16 part(const u32x4 &v, int t)
{
  unsigned w = v[t >> 1];
  unsigned short h = (t & 1) ? (unsigned short)(w >> 16) : (unsigned short)(w & 0xffffu);
  return __builtin_bit_cast(bf16, h);
}

// Stage the 64 x 32 slice of an operand at (row0, k0) into S[row][k] (LDS, row
// stride BK + LPAD).  trans = 0: element (r, k) at P[r ld + k], rows contiguous
// in k: each thread loads 8 k values of one row with one 16-byte load.  trans =
// 1: element at P[k ld + r], contiguous in r: each thread loads 8 rows of one k
// and scatters them.  Rows at or past `rows` and k at or past K are zero.
__device__ __forceinline__ void stage(bf16 (*S)[BK + LPAD], const bf16 *P, int ld, int trans, int row0, int rows,
                                      int k0, int K, int tid)
{
  if (!trans)
  {
    const int r = tid >> 2, c = (tid & 3)*8;
    const bf16 *q = P + (size_t)(row0 + r)*ld + k0 + c;
    if (row0 + r < rows && k0 + c + 8 <= K)
      *reinterpret_cast<u32x4*>(&S[r][c]) = ld8(q);
    else
    {
#pragma unroll
      for (int t = 0; t < 8; t ++) S[r][c + t] = row0 + r < rows && k0 + c + t < K ? q[t] : (bf16)0.0f;
    }
  }
  else
  {
    const int k = tid >> 3, c = (tid & 7)*8;
    const bf16 *q = P + (size_t)(k0 + k)*ld + row0 + c;
    if (k0 + k < K && row0 + c + 8 <= rows)
    {
      const u32x4 v = ld8(q);
#pragma unroll
      for (int t = 0; t < 8; t ++) S[c + t][k] = part(v, t);
    }
    else
    {
#pragma unroll
      for (int t = 0; t < 8; t ++) S[c + t][k] = k0 + k < K && row0 + c + t < rows ? q[t] : (bf16)0.0f;
    }
  }
}

__global__ __launch_bounds__(256) void k_gemm(GemmArgs g)
{
  __shared__ __align__(16) bf16 As[BM][BK + LPAD];
  __shared__ __align__(16) bf16 Bt[BN][BK + LPAD];   // B tile stored transposed: Bt[n][k]
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int m0 = blockIdx.y*BM, n0 = blockIdx.x*BN;
  const int wm = (wave >> 1)*32, wn = (wave & 1)*32;
  const int kb = g.kslice ? blockIdx.z*g.kslice : 0, ke = g.kslice ? min(g.K, kb + g.kslice) : g.K;
  floatx4 acc[2][2];
  for (int i = 0; i < 2; i ++) for (int j = 0; j < 2; j ++) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

  for (int k0 = kb; k0 < ke; k0 += BK)
  {
    // A rows are m (A[m][k] or A[k][m]); B is staged by its n rows: B[n][k]
    // when transB = 1, B[k][n] when transB = 0
    stage(As, g.A, g.lda, g.transA, m0, g.M, k0, ke, tid);
    stage(Bt, g.B, g.ldb, !g.transB, n0, g.N, k0, ke, tid);
    __syncthreads();
    const int fr = lane & 15, fk = 8*(lane >> 4);
    bf16x8 a[2], b[2];
    for (int i = 0; i < 2; i ++) a[i] = *reinterpret_cast<const bf16x8*>(&As[wm + 16*i + fr][fk]);
    for (int j = 0; j < 2; j ++) b[j] = *reinterpret_cast<const bf16x8*>(&Bt[wn + 16*j + fr][fk]);
    for (int i = 0; i < 2; i ++)
      for (int j = 0; j < 2; j ++)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }

  // epilogue: lane holds D[4(l/16)+q][l%16] of each 16x16 tile
  for (int i = 0; i < 2; i ++)
    for (int j = 0; j < 2; j ++)
      for (int q = 0; q < 4; q ++)
      {
        int m = m0 + wm + 16*i + 4*(lane >> 4) + q, n = n0 + wn + 16*j + (lane & 15);
        if (m >= g.M || n >= g.N) continue;
        float v = acc[i][j][q];
        switch (g.epi)
        {
        case EPI_F32: g.C[((size_t)blockIdx.z*g.M + m)*g.ldc + n] = v; break;
        case EPI_BIAS_F32: g.C[(size_t)m*g.ldc + n] = v + g.bias[n]; break;
        case EPI_BIAS_GELU:
        {
          float z = v + g.bias[n];
          g.C[(size_t)m*g.ldc + n] = z;
          g.H[(size_t)m*g.ldh + n] = (bf16)gelu(z);
        }
        break;
        case EPI_GELU_BWD: g.H[(size_t)m*g.ldh + n] = (bf16)(v * dgelu(g.Z[(size_t)m*g.ldz + n])); break;
        case EPI_BF16: g.H[(size_t)m*g.ldh + n] = (bf16)v; break;
        }
      }
}

// ---- model state ----

static const int NF = 18, NI = 32, NH = 128, NO = 3, NOP = 32;

struct Model {
  // FP32 master parameters, layout [in][out] per layer, then biases
  // W1 32x128, b1 128, W2 128x128, b2 128, W3 128x128, b3 128, W4 128x32, b4 32
  size_t off[8], total;
  float *param = nullptr, *grad = nullptr, *m1 = nullptr, *m2 = nullptr;  // Adam moments
  bf16 *wbf = nullptr;      // bf16 copies of W1..W4 (same offsets as the W parts of param)
  float norm[2*NF + NO];   // feature mean, feature std, output scale
  float *dnorm = nullptr;  // device copy of norm
  int have = 0;
  int new_only = 0;        // warm start only contacts without a carried reaction
  float scale = 1.0f;      // trust factor of the prediction
  long long step = 0;
  // activation workspace for a batch
  int cap = 0;
  bf16 *X = nullptr, *H1 = nullptr, *H2 = nullptr, *H3 = nullptr, *dO = nullptr, *dZ = nullptr, *dZ2 = nullptr;
  float *Z1 = nullptr, *Z2 = nullptr, *Z3 = nullptr, *O = nullptr, *Y = nullptr, *loss = nullptr;
  float *part = nullptr;   // split-K partials of the weight gradients: ceil(cap / KSLICE) x 128 x 128
  // device-resident training set
  float *tX = nullptr, *tY = nullptr;
  long long tn = 0, tcap = 0;
  int *idx = nullptr;
};

static void layout(Model &m)
{
  size_t o = 0;
  m.off[0] = o; o += (size_t)NI*NH;  m.off[1] = o; o += NH;
  m.off[2] = o; o += (size_t)NH*NH;  m.off[3] = o; o += NH;
  m.off[4] = o; o += (size_t)NH*NH;  m.off[5] = o; o += NH;
  m.off[6] = o; o += (size_t)NH*NOP; m.off[7] = o; o += NOP;
  m.total = o;
}

// out[i] = sum over the S split-K partials part[z n + i], in z order (deterministic)
__global__ void k_sum_parts(int S, int n, const float *part, float *out)
{
  int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i >= n) return;
  float v = 0.0f;
  for (int z = 0; z < S; z ++) v += part[(size_t)z*n + i];
  out[i] = v;
}

__global__ void k_to_bf16(size_t n, const float *src, bf16 *dst)
{
  size_t i = (size_t)blockIdx.x*blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (bf16)src[i];
}

// MSE loss on the 3 used outputs; dO = d loss / d out (bf16, padded columns zero)
__global__ void k_loss(int B, const float *O, const float *Y, bf16 *dO, float *loss_part)
{
  __shared__ float red[256];
  int i = blockIdx.x*blockDim.x + threadIdx.x;
  float l = 0.f;
  if (i < B)
  {
    for (int c = 0; c < NOP; c ++)
    {
      float g = 0.f;
      if (c < NO)
      {
        float d = O[(size_t)i*NOP + c] - Y[(size_t)i*NO + c];
        l += d*d;
        g = 2.f*d / (float)(B*NO);
      }
      dO[(size_t)i*NOP + c] = (bf16)g;
    }
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1)
  {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

// bias gradient: column sums of a bf16 B x N matrix
__global__ void k_colsum(int B, int N, const bf16 *D, float *g)
{
  int n = blockIdx.x*blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0.f;
  for (int i = 0; i < B; i ++) s += (float)D[(size_t)i*N + n];
  g[n] = s;
}

__global__ void k_adam(size_t n, float *p, const float *g, float *m1, float *m2, float lr, float b1, float b2,
                       float c1, float c2, float eps)
{
  size_t i = (size_t)blockIdx.x*blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gi = g[i];
  float a = b1*m1[i] + (1.f - b1)*gi, v = b2*m2[i] + (1.f - b2)*gi*gi;
  m1[i] = a; m2[i] = v;
  p[i] -= lr * (a/c1) / (sqrtf(v/c2) + eps);
}

// minibatch gather from the resident training set: normalised features -> bf16 X (B x 32), targets Y
__global__ void k_gather(int B, const int *idx, const float *tX, const float *tY, const float *norm, bf16 *X, float *Y)
{
  int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i >= B) return;
  int s = idx[i];
  for (int f = 0; f < NI; f ++)
  {
    float v = 0.f;
    if (f < NF) v = (tX[(size_t)s*NF + f] - norm[f]) / norm[NF + f];
    X[(size_t)i*NI + f] = (bf16)v;
  }
  for (int c = 0; c < NO; c ++) Y[(size_t)i*NO + c] = tY[(size_t)s*NO + c] / norm[2*NF + c];
}

// ---- fused inference of the warm start ----
//
// The whole network for tiles of 64 rows in one persistent launch: normalised
// features into LDS as bf16, three hidden layers through LDS (ping-pong
// buffers of 64 x 128 bf16), the output layer and the reaction shift, without
// any activation in HBM.  Wave w owns columns 32w .. 32w + 31 of every hidden
// layer and keeps its B fragments of W1, W2, W3 (and the output layer's first
// 16 columns) in registers for the whole launch, loaded once per workgroup.
// Every output element is formed as k_gemm forms it (the same 16x16x32 bf16
// products in the same k order, the same lane placement, bias then GELU), so
// the result is that of the layer-by-layer forward pass.
struct InferArgs {
  int n;
  const float *feat;       // n x 18, rows in the order of R
  const float *norm;       // feature mean (18), std (18), output scale (3)
  const bf16 *W1, *W2, *W3, *W4;   // [in][out]: 32 x 128, 128 x 128, 128 x 128, 128 x 32
  const float *b1, *b2, *b3, *b4;
  double *R;               // 4 per row
  int new_only;
  float scale;
};

static const int IT_M = 64;   // rows per tile

// B fragment of a [K][ldw] weight matrix: column n, k = k0 .. k0 + 7
__device__ __forceinline__ bf16x8 wfrag(const bf16 *W, int ldw, int n, int k0)
{
  bf16x8 f;
#pragma unroll
  for (int t = 0; t < 8; t ++) f[t] = W[(size_t)(k0 + t)*ldw + n];
  return f;
}

// one hidden layer of the tile: Out = bf16(GELU(In Wt + b)) for this wave's
// 32 columns, K = 128 in four chunks of 32 (the k order of k_gemm)
__device__ __forceinline__ void infer_hidden(const bf16 (*In)[NH + LPAD], bf16 (*Out)[NH + LPAD], const bf16x8 (&w)[2][4],
                                             const float (&b)[2], int wave, int lane)
{
  const int fr = lane & 15, fk = 8*(lane >> 4);
  floatx4 acc[4][2];
#pragma unroll
  for (int i = 0; i < 4; i ++)
#pragma unroll
    for (int j = 0; j < 2; j ++) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int c = 0; c < 4; c ++)
#pragma unroll
    for (int i = 0; i < 4; i ++)
    {
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(&In[16*i + fr][32*c + fk]);
#pragma unroll
      for (int j = 0; j < 2; j ++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, w[j][c], acc[i][j], 0, 0, 0);
    }
#pragma unroll
  for (int i = 0; i < 4; i ++)
#pragma unroll
    for (int j = 0; j < 2; j ++)
#pragma unroll
      for (int q = 0; q < 4; q ++)
        Out[16*i + 4*(lane >> 4) + q][32*wave + 16*j + fr] = (bf16)gelu(acc[i][j][q] + b[j]);
}

__global__ __launch_bounds__(256) void k_mlp_infer(InferArgs a)
{
  __shared__ __align__(16) bf16 Xs[IT_M][NI + LPAD];
  __shared__ __align__(16) bf16 Ha[IT_M][NH + LPAD];
  __shared__ __align__(16) bf16 Hb[IT_M][NH + LPAD];
  __shared__ float Os[IT_M][NO];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, fr = lane & 15, fk = 8*(lane >> 4);
  bf16x8 w1[2], w2[2][4], w3[2][4], w4[4];
  float b1[2], b2[2], b3[2];
#pragma unroll
  for (int j = 0; j < 2; j ++)
  {
    const int n = 32*wave + 16*j + fr;
    w1[j] = wfrag(a.W1, NH, n, fk);
#pragma unroll
    for (int c = 0; c < 4; c ++) { w2[j][c] = wfrag(a.W2, NH, n, 32*c + fk); w3[j][c] = wfrag(a.W3, NH, n, 32*c + fk); }
    b1[j] = a.b1[n]; b2[j] = a.b2[n]; b3[j] = a.b3[n];
  }
#pragma unroll
  for (int c = 0; c < 4; c ++) w4[c] = wfrag(a.W4, NOP, fr, 32*c + fk);
  const float b4 = a.b4[fr];

  for (int row0 = blockIdx.x*IT_M; row0 < a.n; row0 += gridDim.x*IT_M)
  {
    // normalised features (k_gather): thread t takes row t/4, features 8(t%4) .. +7
    {
      const int r = tid >> 2, f0 = 8*(tid & 3);
#pragma unroll
      for (int t = 0; t < 8; t ++)
      {
        const int f = f0 + t;
        float v = 0.f;
        if (row0 + r < a.n && f < NF) v = (a.feat[(size_t)(row0 + r)*NF + f] - a.norm[f]) / a.norm[NF + f];
        Xs[r][f] = (bf16)v;
      }
    }
    __syncthreads();
    // layer 1 (K = 32: one chunk)
    {
      floatx4 acc[4][2];
#pragma unroll
      for (int i = 0; i < 4; i ++)
      {
        const bf16x8 x = *reinterpret_cast<const bf16x8*>(&Xs[16*i + fr][fk]);
#pragma unroll
        for (int j = 0; j < 2; j ++)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x, w1[j], floatx4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
      }
#pragma unroll
      for (int i = 0; i < 4; i ++)
#pragma unroll
        for (int j = 0; j < 2; j ++)
#pragma unroll
          for (int q = 0; q < 4; q ++)
            Ha[16*i + 4*(lane >> 4) + q][32*wave + 16*j + fr] = (bf16)gelu(acc[i][j][q] + b1[j]);
    }
    __syncthreads();
    infer_hidden(Ha, Hb, w2, b2, wave, lane);
    __syncthreads();
    infer_hidden(Hb, Ha, w3, b3, wave, lane);
    __syncthreads();
    // output layer: wave w takes rows 16w .. 16w + 15, the first 16 output columns
    {
      floatx4 acc = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int c = 0; c < 4; c ++)
      {
        const bf16x8 h = *reinterpret_cast<const bf16x8*>(&Ha[16*wave + fr][32*c + fk]);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(h, w4[c], acc, 0, 0, 0);
      }
      if (fr < NO)
#pragma unroll
        for (int q = 0; q < 4; q ++) Os[16*wave + 4*(lane >> 4) + q][fr] = acc[q] + b4;
    }
    __syncthreads();
    // the shift of R: R0 = R_prev + scale * out * output scale; with new_only,
    // rows that carry a reaction keep it
    if (tid < IT_M && row0 + tid < a.n)
    {
      double *r = a.R + 4*(size_t)(row0 + tid);
      if (!(a.new_only && (r[0] != 0.0 || r[1] != 0.0 || r[2] != 0.0)))
        for (int c = 0; c < NO; c ++) r[c] += (double)(a.scale * Os[tid][c] * a.norm[2*NF + c]);
    }
    __syncthreads();   // Xs, Ha and Os are rewritten by the next tile
  }
}

} // namespace sfgs_mlp

using namespace sfgs_mlp;

#define HIPCHK(call) do { hipError_t e_ = (call); if (e_ != hipSuccess) { \
    ctx->err = std::string(#call) + ": " + hipGetErrorString(e_); return SFGS_E_HIP; } } while (0)

static Model* model_of(sfgs_ctx *ctx)
{
  if (!ctx->mlp)
  {
    Model *m = new Model();
    layout(*m);
    ctx->mlp = m;
  }
  return (Model*)ctx->mlp;
}

void sfgs_mlp_release(sfgs_ctx *ctx)
{
  Model *m = (Model*)ctx->mlp;
  if (!m) return;
  void *ptrs[] = {m->param, m->grad, m->m1, m->m2, m->wbf, m->dnorm, m->part, m->X, m->H1, m->H2, m->H3, m->dO, m->dZ, m->dZ2,
                  m->Z1, m->Z2, m->Z3, m->O, m->Y, m->loss, m->tX, m->tY, m->idx};
  for (void *p : ptrs) if (p) (void)hipFree(p);
  delete m;
  ctx->mlp = nullptr;
}

static const int KSLICE = 512;   // K per workgroup of a split-K weight-gradient product

static int gemm(sfgs_ctx *ctx, const GemmArgs &g)
{
  dim3 grid((g.N + BN - 1)/BN, (g.M + BM - 1)/BM, g.kslice ? (g.K + g.kslice - 1)/g.kslice : 1);
  hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, ctx->stream, g);
  return 0;
}

// a weight gradient C = A' B (small M x N, K = the batch): K split over
// workgroups in slices of KSLICE, the partials summed in slice order into out
static int gemm_wgrad(sfgs_ctx *ctx, Model *m, GemmArgs g, float *out)
{
  const int S = (g.K + KSLICE - 1)/KSLICE;
  g.ldc = g.N;
  if (S <= 1) { g.C = out; return gemm(ctx, g); }
  g.kslice = KSLICE;
  g.C = m->part;
  gemm(ctx, g);
  const int n = g.M*g.N;
  hipLaunchKernelGGL(k_sum_parts, dim3((n + 255)/256), dim3(256), 0, ctx->stream, S, n, m->part, out);
  return 0;
}

static GemmArgs gargs(int M, int N, int K, const bf16 *A, int lda, int tA, const bf16 *B, int ldb, int tB, int epi)
{
  GemmArgs g;
  std::memset(&g, 0, sizeof(g));
  g.M = M; g.N = N; g.K = K; g.A = A; g.lda = lda; g.transA = tA; g.B = B; g.ldb = ldb; g.transB = tB; g.epi = epi;
  return g;
}

static int ensure_batch(sfgs_ctx *ctx, Model *m, int B)
{
  if (B <= m->cap) return 0;
  (void)hipStreamSynchronize(ctx->stream);
  void **bufs[] = {(void**)&m->X, (void**)&m->H1, (void**)&m->H2, (void**)&m->H3, (void**)&m->dO, (void**)&m->dZ,
                   (void**)&m->dZ2, (void**)&m->Z1, (void**)&m->Z2, (void**)&m->Z3, (void**)&m->O, (void**)&m->Y,
                   (void**)&m->loss};
  size_t sz[] = {sizeof(bf16)*NI, sizeof(bf16)*NH, sizeof(bf16)*NH, sizeof(bf16)*NH, sizeof(bf16)*NOP, sizeof(bf16)*NH,
                 sizeof(bf16)*NH, sizeof(float)*NH, sizeof(float)*NH, sizeof(float)*NH, sizeof(float)*NOP,
                 sizeof(float)*NO, sizeof(float)};
  int cap = std::max(B, 1024);
  for (int k = 0; k < 13; k ++)
  {
    if (*bufs[k]) (void)hipFree(*bufs[k]);
    *bufs[k] = nullptr;
    HIPCHK(hipMalloc(bufs[k], sz[k]*(size_t)cap));
  }
  if (m->part) (void)hipFree(m->part);
  m->part = nullptr;
  HIPCHK(hipMalloc(&m->part, sizeof(float)*(size_t)NH*NH*(size_t)((cap + KSLICE - 1)/KSLICE)));
  m->cap = cap;
  return 0;
}

// forward pass of B rows of m->X (bf16, B x 32) into m->O (FP32, B x 32)
static int forward(sfgs_ctx *ctx, Model *m, int B)
{
  const bf16 *W1 = m->wbf + m->off[0], *W2 = m->wbf + m->off[2], *W3 = m->wbf + m->off[4], *W4 = m->wbf + m->off[6];
  const float *b1 = m->param + m->off[1], *b2 = m->param + m->off[3], *b3 = m->param + m->off[5], *b4 = m->param + m->off[7];
  GemmArgs g = gargs(B, NH, NI, m->X, NI, 0, W1, NH, 0, EPI_BIAS_GELU);
  g.C = m->Z1; g.ldc = NH; g.H = m->H1; g.ldh = NH; g.bias = b1;
  gemm(ctx, g);
  g = gargs(B, NH, NH, m->H1, NH, 0, W2, NH, 0, EPI_BIAS_GELU);
  g.C = m->Z2; g.ldc = NH; g.H = m->H2; g.ldh = NH; g.bias = b2;
  gemm(ctx, g);
  g = gargs(B, NH, NH, m->H2, NH, 0, W3, NH, 0, EPI_BIAS_GELU);
  g.C = m->Z3; g.ldc = NH; g.H = m->H3; g.ldh = NH; g.bias = b3;
  gemm(ctx, g);
  g = gargs(B, NOP, NH, m->H3, NH, 0, W4, NOP, 0, EPI_BIAS_F32);
  g.C = m->O; g.ldc = NOP; g.bias = b4;
  gemm(ctx, g);
  return 0;
}

// backward pass from m->dO (bf16) into m->grad (FP32); forward activations must be current
static int backward(sfgs_ctx *ctx, Model *m, int B)
{
  const bf16 *W2 = m->wbf + m->off[2], *W3 = m->wbf + m->off[4], *W4 = m->wbf + m->off[6];
  float *G = m->grad;
  // layer 4
  GemmArgs g = gargs(NH, NOP, B, m->H3, NH, 1, m->dO, NOP, 0, EPI_F32);     // dW4 = H3' dO
  gemm_wgrad(ctx, m, g, G + m->off[6]);
  hipLaunchKernelGGL(k_colsum, dim3(1), dim3(64), 0, ctx->stream, B, NOP, m->dO, G + m->off[7]);
  g = gargs(B, NH, NOP, m->dO, NOP, 0, W4, NOP, 1, EPI_GELU_BWD);          // dZ3 = (dO W4') * gelu'(Z3)
  g.H = m->dZ; g.ldh = NH; g.Z = m->Z3; g.ldz = NH;
  gemm(ctx, g);
  // layer 3
  g = gargs(NH, NH, B, m->H2, NH, 1, m->dZ, NH, 0, EPI_F32);               // dW3 = H2' dZ3
  gemm_wgrad(ctx, m, g, G + m->off[4]);
  hipLaunchKernelGGL(k_colsum, dim3(2), dim3(64), 0, ctx->stream, B, NH, m->dZ, G + m->off[5]);
  g = gargs(B, NH, NH, m->dZ, NH, 0, W3, NH, 1, EPI_GELU_BWD);             // dZ2 = (dZ3 W3') * gelu'(Z2)
  g.H = m->dZ2; g.ldh = NH; g.Z = m->Z2; g.ldz = NH;
  gemm(ctx, g);
  // layer 2
  g = gargs(NH, NH, B, m->H1, NH, 1, m->dZ2, NH, 0, EPI_F32);              // dW2 = H1' dZ2
  gemm_wgrad(ctx, m, g, G + m->off[2]);
  hipLaunchKernelGGL(k_colsum, dim3(2), dim3(64), 0, ctx->stream, B, NH, m->dZ2, G + m->off[3]);
  g = gargs(B, NH, NH, m->dZ2, NH, 0, W2, NH, 1, EPI_GELU_BWD);            // dZ1 = (dZ2 W2') * gelu'(Z1)
  g.H = m->dZ; g.ldh = NH; g.Z = m->Z1; g.ldz = NH;
  gemm(ctx, g);
  // layer 1
  g = gargs(NI, NH, B, m->X, NI, 1, m->dZ, NH, 0, EPI_F32);                // dW1 = X' dZ1
  gemm_wgrad(ctx, m, g, G + m->off[0]);
  hipLaunchKernelGGL(k_colsum, dim3(2), dim3(64), 0, ctx->stream, B, NH, m->dZ, G + m->off[1]);
  return 0;
}

static int refresh_bf16(sfgs_ctx *ctx, Model *m)
{
  size_t n = m->total;
  hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((n + 255)/256)), dim3(256), 0, ctx->stream, n, m->param, m->wbf);
  return 0;
}

// apply the model to device features (n x 18, row order of R) and shift R by
// the prediction: one fused launch (k_mlp_infer), no workspace
int sfgs_mlp_apply(sfgs_ctx *ctx, const float *feat, int n, double *R)
{
  Model *m = model_of(ctx);
  if (!m->have) { ctx->err = "no MLP parameters"; return SFGS_E_STATE; }
  if (n == 0) return 0;
  InferArgs a;
  a.n = n; a.feat = feat; a.norm = m->dnorm;
  a.W1 = m->wbf + m->off[0]; a.W2 = m->wbf + m->off[2]; a.W3 = m->wbf + m->off[4]; a.W4 = m->wbf + m->off[6];
  a.b1 = m->param + m->off[1]; a.b2 = m->param + m->off[3]; a.b3 = m->param + m->off[5]; a.b4 = m->param + m->off[7];
  a.R = R; a.new_only = m->new_only; a.scale = m->scale;
  const int tiles = (n + IT_M - 1)/IT_M;
  hipLaunchKernelGGL(k_mlp_infer, dim3(std::min(tiles, 512)), dim3(256), 0, ctx->stream, a);   // two per CU resident
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" {

int SFGS_Mlp_Set(sfgs_ctx *ctx, const float *params, long long nparams, const float *norm)
{
  if (!ctx || !params || !norm) return SFGS_E_ARG;
  Model *m = model_of(ctx);
  if ((size_t)nparams != m->total) { ctx->err = "MLP parameter count must be " + std::to_string(m->total); return SFGS_E_ARG; }
  HIPCHK(hipSetDevice(ctx->device));
  if (!m->param)
  {
    HIPCHK(hipMalloc(&m->param, sizeof(float)*m->total));
    HIPCHK(hipMalloc(&m->grad, sizeof(float)*m->total));
    HIPCHK(hipMalloc(&m->m1, sizeof(float)*m->total));
    HIPCHK(hipMalloc(&m->m2, sizeof(float)*m->total));
    HIPCHK(hipMalloc(&m->wbf, sizeof(bf16)*m->total));
    HIPCHK(hipMalloc(&m->dnorm, sizeof(m->norm)));
  }
  HIPCHK(hipMemcpyAsync(m->param, params, sizeof(float)*m->total, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemsetAsync(m->m1, 0, sizeof(float)*m->total, ctx->stream));
  HIPCHK(hipMemsetAsync(m->m2, 0, sizeof(float)*m->total, ctx->stream));
  std::memcpy(m->norm, norm, sizeof(m->norm));
  for (int f = 0; f < NF; f ++) if (!(m->norm[NF + f] > 0.f)) m->norm[NF + f] = 1.f;
  for (int c = 0; c < NO; c ++) if (!(m->norm[2*NF + c] > 0.f)) m->norm[2*NF + c] = 1.f;
  HIPCHK(hipMemcpyAsync(m->dnorm, m->norm, sizeof(m->norm), hipMemcpyHostToDevice, ctx->stream));
  refresh_bf16(ctx, m);
  HIPCHK(hipStreamSynchronize(ctx->stream));
  m->have = 1;
  m->step = 0;
  return 0;
}

int SFGS_Warm_Start_Mode(sfgs_ctx *ctx, int new_only, float scale)
{
  if (!ctx || (new_only != 0 && new_only != 1) || !(scale >= 0.0f)) return SFGS_E_ARG;
  Model *m = model_of(ctx);
  m->new_only = new_only;
  m->scale = scale;
  return 0;
}

int SFGS_Mlp_Get(sfgs_ctx *ctx, float *params, long long *nparams)
{
  if (!ctx) return SFGS_E_ARG;
  Model *m = model_of(ctx);
  if (nparams) *nparams = (long long)m->total;
  if (!params) return 0;
  if (!m->have) { ctx->err = "no MLP parameters"; return SFGS_E_STATE; }
  HIPCHK(hipMemcpyAsync(params, m->param, sizeof(float)*m->total, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return 0;
}

// forward pass on host features (n x 18, raw): out (n x 3) in target units (R - R_prev)
int SFGS_Mlp_Forward(sfgs_ctx *ctx, const float *features, int n, float *out)
{
  if (!ctx || n < 0 || (n > 0 && (!features || !out))) return SFGS_E_ARG;
  Model *m = model_of(ctx);
  if (!m->have) { ctx->err = "no MLP parameters"; return SFGS_E_STATE; }
  if (n == 0) return 0;
  HIPCHK(hipSetDevice(ctx->device));
  if (ensure_batch(ctx, m, n)) return SFGS_E_HIP;
  float *tX = nullptr;
  int *idx = nullptr;
  const float *dn = m->dnorm;
  HIPCHK(hipMalloc(&tX, sizeof(float)*NF*(size_t)n));
  HIPCHK(hipMalloc(&idx, sizeof(int)*(size_t)n));
  std::vector<int> hid((size_t)n);
  for (int i = 0; i < n; i ++) hid[i] = i;
  HIPCHK(hipMemcpyAsync(tX, features, sizeof(float)*NF*(size_t)n, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(idx, hid.data(), sizeof(int)*(size_t)n, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_gather, dim3((n + 255)/256), dim3(256), 0, ctx->stream, n, idx, tX, tX, dn, m->X, m->Y);
  forward(ctx, m, n);
  std::vector<float> o((size_t)n*NOP);
  HIPCHK(hipMemcpyAsync(o.data(), m->O, sizeof(float)*NOP*(size_t)n, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipGetLastError());
  (void)hipFree(tX); (void)hipFree(idx);
  for (int i = 0; i < n; i ++)
    for (int c = 0; c < NO; c ++) out[(size_t)i*NO + c] = o[(size_t)i*NOP + c] * m->norm[2*NF + c];
  return 0;
}

// loss and gradient of one batch of host samples (features n x 18 raw, targets n x 3 in
// target units): grad receives the FP32 gradient of the mean squared error in
// normalised units with respect to every parameter (layout of SFGS_Mlp_Set)
int SFGS_Mlp_Grad(sfgs_ctx *ctx, const float *features, const float *targets, int n, float *grad, float *loss)
{
  if (!ctx || n <= 0 || !features || !targets) return SFGS_E_ARG;
  Model *m = model_of(ctx);
  if (!m->have) { ctx->err = "no MLP parameters"; return SFGS_E_STATE; }
  HIPCHK(hipSetDevice(ctx->device));
  if (ensure_batch(ctx, m, n)) return SFGS_E_HIP;
  float *tX = nullptr, *tY = nullptr;
  int *idx = nullptr;
  const float *dn = m->dnorm;
  HIPCHK(hipMalloc(&tX, sizeof(float)*NF*(size_t)n));
  HIPCHK(hipMalloc(&tY, sizeof(float)*NO*(size_t)n));
  HIPCHK(hipMalloc(&idx, sizeof(int)*(size_t)n));
  std::vector<int> hid((size_t)n);
  for (int i = 0; i < n; i ++) hid[i] = i;
  HIPCHK(hipMemcpyAsync(tX, features, sizeof(float)*NF*(size_t)n, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(tY, targets, sizeof(float)*NO*(size_t)n, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(idx, hid.data(), sizeof(int)*(size_t)n, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_gather, dim3((n + 255)/256), dim3(256), 0, ctx->stream, n, idx, tX, tY, dn, m->X, m->Y);
  forward(ctx, m, n);
  int lb = (n + 255)/256;
  hipLaunchKernelGGL(k_loss, dim3(lb), dim3(256), 0, ctx->stream, n, m->O, m->Y, m->dO, m->loss);
  backward(ctx, m, n);
  std::vector<float> lp((size_t)lb);
  HIPCHK(hipMemcpyAsync(lp.data(), m->loss, sizeof(float)*(size_t)lb, hipMemcpyDeviceToHost, ctx->stream));
  if (grad) HIPCHK(hipMemcpyAsync(grad, m->grad, sizeof(float)*m->total, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipGetLastError());
  (void)hipFree(tX); (void)hipFree(tY); (void)hipFree(idx);
  if (loss)
  {
    double s = 0.0;
    for (float v : lp) s += v;
    *loss = (float)(s / ((double)n*NO));
  }
  return 0;
}

// Adam training on a device-resident copy of the samples: epochs over random
// minibatches (seeded permutation on the host), loss per epoch into loss_hist
int SFGS_Mlp_Train(sfgs_ctx *ctx, const float *features, const float *targets, long long n, int epochs, int batch,
                   float lr, unsigned long long seed, float *loss_hist)
{
  if (!ctx || n <= 0 || epochs <= 0 || batch <= 0 || !features || !targets) return SFGS_E_ARG;
  Model *m = model_of(ctx);
  if (!m->have) { ctx->err = "no MLP parameters"; return SFGS_E_STATE; }
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  if (batch > n) batch = (int)n;
  if (ensure_batch(ctx, m, batch)) return SFGS_E_HIP;
  if (n > m->tcap)
  {
    if (m->tX) (void)hipFree(m->tX);
    if (m->tY) (void)hipFree(m->tY);
    if (m->idx) (void)hipFree(m->idx);
    m->tX = m->tY = nullptr; m->idx = nullptr;
    HIPCHK(hipMalloc(&m->tX, sizeof(float)*NF*(size_t)n));
    HIPCHK(hipMalloc(&m->tY, sizeof(float)*NO*(size_t)n));
    HIPCHK(hipMalloc(&m->idx, sizeof(int)*(size_t)n));
    m->tcap = n;
  }
  const float *dn = m->dnorm;
  HIPCHK(hipMemcpyAsync(m->tX, features, sizeof(float)*NF*(size_t)n, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(m->tY, targets, sizeof(float)*NO*(size_t)n, hipMemcpyHostToDevice, st));
  std::vector<int> perm((size_t)n);
  unsigned long long s = seed;
  const int lbmax = (batch + 255)/256;
  std::vector<float> lp((size_t)lbmax);
  const float b1 = 0.9f, b2 = 0.999f;
  for (int e = 0; e < epochs; e ++)
  {
    for (long long i = 0; i < n; i ++) perm[i] = (int)i;
    for (long long i = n - 1; i > 0; i --)   // Fisher-Yates with splitmix64
    {
      s += 0x9e3779b97f4a7c15ULL;
      unsigned long long z = s;
      z = (z ^ (z >> 30))*0xbf58476d1ce4e5b9ULL; z = (z ^ (z >> 27))*0x94d049bb133111ebULL; z ^= z >> 31;
      long long j = (long long)(z % (unsigned long long)(i + 1));
      std::swap(perm[i], perm[j]);
    }
    HIPCHK(hipMemcpyAsync(m->idx, perm.data(), sizeof(int)*(size_t)n, hipMemcpyHostToDevice, st));
    double esum = 0.0;
    long long cnt = 0;
    for (long long b0 = 0; b0 + batch <= n; b0 += batch)
    {
      hipLaunchKernelGGL(k_gather, dim3((batch + 255)/256), dim3(256), 0, st, batch, m->idx + b0, m->tX, m->tY, dn,
                         m->X, m->Y);
      forward(ctx, m, batch);
      hipLaunchKernelGGL(k_loss, dim3(lbmax), dim3(256), 0, st, batch, m->O, m->Y, m->dO, m->loss);
      backward(ctx, m, batch);
      m->step ++;
      float c1 = 1.f - powf(b1, (float)m->step), c2 = 1.f - powf(b2, (float)m->step);
      hipLaunchKernelGGL(k_adam, dim3((unsigned)((m->total + 255)/256)), dim3(256), 0, st, m->total, m->param, m->grad,
                         m->m1, m->m2, lr, b1, b2, c1, c2, 1e-8f);
      refresh_bf16(ctx, m);
      HIPCHK(hipMemcpyAsync(lp.data(), m->loss, sizeof(float)*(size_t)lbmax, hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      for (float v : lp) esum += v;
      cnt += (long long)batch*NO;
    }
    if (loss_hist) loss_hist[e] = cnt ? (float)(esum / (double)cnt) : 0.f;
  }
  HIPCHK(hipGetLastError());
  return 0;
}

} // extern "C"
