// sfgs_newton.hip -- the NEWTON_SOLVER (nts.c) on the assembled local dynamics
// of a context: a projected semismooth Newton method over all contacts at once,
// each step a linear solve by flexible GMRES (or a per-row "diagonalised"
// step), then the Gauss-Seidel solve as the fallback of an unconverged run.
//
// Vectors of the Newton step are 3 doubles per row in the colour-major row
// order of the store, so the W products reuse the SELL-64 slices the sweeps
// stream.  The kernels:
//   k_nt_lin     per row: C(U, R), X = dC/dU, Y = dC/dR of the smoothed cone
//                residual (scf.c:28-133), b = -C, T = X W_ii + Y
//                (nts.c:1072-1081); GMRES: P = (T + delta I)^-1
//                (nts.c:1095-1103); DIAG: T d = b, dr = (1 - theta) dr +
//                theta d, R = proj(R + dr) (nts.c:1084-1093)
//   k_nt_matvec  y = beta y + alpha (X (W x) + Y x) + alpha delta x: the
//                operator of the linear step (Matvec, nts.c:584-652), W x
//                over the slices with the row's lanes summed by shuffles
//   k_nt_dot     deterministic per-block partial sums of x'y (summed on the
//                host in block order)
//   k_nt_axpy, k_nt_scale, k_nt_precond, k_nt_update  the vector work of
//                flexible GMRES (ext/krylov/flexgmres.c) and R = proj(R + dr)
// The GMRES driver runs on the host: its Hessenberg matrix is at most 21 x 20
// and every dot product is read back, as the reference's InnerProd returns a
// value.  All of this is HBM-bound vector work (a W product streams the same
// bytes as a sweep); the sweeps remain the solver of the hot path.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include <cfloat>
#include <vector>
#include <algorithm>
#include "sfgs_ctx.hpp"
#include "sfgs_math.hpp"
#include "sfgs_host.hpp"

using namespace sfgs;

namespace {

#define NT_DISPATCH_G(G, CALL) \
  switch (G) { \
  case 1: { constexpr int GG = 1; CALL; } break; \
  case 2: { constexpr int GG = 2; CALL; } break; \
  case 8: { constexpr int GG = 8; CALL; } break; \
  default: { constexpr int GG = 4; CALL; } break; }

constexpr int NT_KDIM = 20;        // hypre_FlexGMRESCreate default k_dim (flexgmres.c:91)
constexpr int NT_DOT_BLOCKS = 256; // partials of one dot product

struct NtRows {
  int n;
  const double *R;      // 4 per row
  const double *U;      // SoA planes
  const double *Wd;     // SoA planes
  const double *fri, *coh, *shift;
  double *X, *Y, *P;    // SoA planes, 9 per row
  double *b, *dr;       // 3 per row
  int *singular;
};

__global__ __launch_bounds__(TPB) void k_nt_lin(NtRows a, int diag, double omega, double delta, double theta,
                                                 double *Rw)
{
  const int p = blockIdx.x*blockDim.x + threadIdx.x;
  if (p >= a.n) return;
  const int n = a.n;
  double R[3], U[3], W[9], C[3], X[9], Y[9], T[9];
#pragma unroll
  for (int k = 0; k < 3; k ++) { R[k] = a.R[4*(size_t)p + k]; U[k] = a.U[(size_t)k*n + p]; }
#pragma unroll
  for (int e = 0; e < 9; e ++) W[e] = a.Wd[(size_t)e*n + p];
  const double fri = a.fri[p], coh = a.coh[p];
  scf_linearize_smooth(fri, coh, a.shift[p], omega, U, R, C, X, Y);
  double b[3] = {-C[0], -C[1], -C[2]};
#pragma unroll
  for (int r = 0; r < 3; r ++)
#pragma unroll
    for (int c = 0; c < 3; c ++)
      T[r + 3*c] = (X[r]*W[3*c] + X[r + 3]*W[3*c + 1] + X[r + 6]*W[3*c + 2]) + Y[r + 3*c];
  if (diag)
  {
    if (solve3(T, b)) { atomicAdd(a.singular, 1); b[0] = b[1] = b[2] = 0.0; }
    const double gamma = 1.0 - theta;
    double d[3];
#pragma unroll
    for (int k = 0; k < 3; k ++)
    {
      d[k] = gamma*a.dr[3*(size_t)p + k] + theta*b[k];
      a.dr[3*(size_t)p + k] = d[k];
      R[k] += d[k];
    }
    double Rp[3];
    scf_project(fri, coh, R, Rp);
#pragma unroll
    for (int k = 0; k < 3; k ++) Rw[4*(size_t)p + k] = Rp[k];
    return;
  }
  T[0] += delta; T[4] += delta; T[8] += delta;
  double Pm[9];
  if (inverse3(T, Pm)) atomicAdd(a.singular, 1);
#pragma unroll
  for (int e = 0; e < 9; e ++)
  {
    a.X[(size_t)e*n + p] = X[e];
    a.Y[(size_t)e*n + p] = Y[e];
    a.P[(size_t)e*n + p] = Pm[e];
  }
#pragma unroll
  for (int k = 0; k < 3; k ++) a.b[3*(size_t)p + k] = b[k];
}

struct NtMat {
  const int *row0, *nrows, *rowlen;
  const long long *off;
  const double *W;
  const int *wcol;
  const double *Wd, *X, *Y;
  int n;
  double delta;
};

template <int G>
__global__ __launch_bounds__(TPB) void k_nt_matvec(NtMat a, int nslices, double alpha, const double *x, double beta,
                                                    double *y)
{
  const int s = blockIdx.x*(TPB/SLICE) + (threadIdx.x >> 6);
  if (s >= nslices) return;   // uniform over the wavefront
  const int l = threadIdx.x & 63, r = l / G, g = l % G, n = a.n;
  const bool act = r < a.nrows[s];
  const long long off = a.off[s];
  const int p = a.row0[s] + r, len = act ? a.rowlen[p] : 0;
  double z0 = 0.0, z1 = 0.0, z2 = 0.0;
  const double *Wb = a.W + 9*off + l;
  const int *cb = a.wcol + off + l;
  for (int k = g, m = 0; k < len; k += G, m ++)
  {
    const int j = cb[(size_t)m*SLICE];
    const double *w = Wb + 9*(size_t)m*SLICE;
    const double x0 = x[3*(size_t)j], x1 = x[3*(size_t)j + 1], x2 = x[3*(size_t)j + 2];
    z0 += w[0]*x0 + w[3*SLICE]*x1 + w[6*SLICE]*x2;
    z1 += w[SLICE]*x0 + w[4*SLICE]*x1 + w[7*SLICE]*x2;
    z2 += w[2*SLICE]*x0 + w[5*SLICE]*x1 + w[8*SLICE]*x2;
  }
  z0 = gsum<G>(z0); z1 = gsum<G>(z1); z2 = gsum<G>(z2);
  if (!act || g != 0) return;
  const double xp[3] = {x[3*(size_t)p], x[3*(size_t)p + 1], x[3*(size_t)p + 2]};
  double Wi[9], X[9], Y[9];
#pragma unroll
  for (int e = 0; e < 9; e ++)
  {
    Wi[e] = a.Wd[(size_t)e*n + p];
    X[e] = a.X[(size_t)e*n + p];
    Y[e] = a.Y[(size_t)e*n + p];
  }
  const double z[3] = {Wi[0]*xp[0] + Wi[3]*xp[1] + Wi[6]*xp[2] + z0,
                       Wi[1]*xp[0] + Wi[4]*xp[1] + Wi[7]*xp[2] + z1,
                       Wi[2]*xp[0] + Wi[5]*xp[1] + Wi[8]*xp[2] + z2};
  const double ad = alpha*a.delta;
#pragma unroll
  for (int k = 0; k < 3; k ++)
  {
    const double u = (X[k]*z[0] + X[k + 3]*z[1] + X[k + 6]*z[2]) + (Y[k]*xp[0] + Y[k + 3]*xp[1] + Y[k + 6]*xp[2]);
    double q = beta*y[3*(size_t)p + k];
    q = q + alpha*u;
    y[3*(size_t)p + k] = q + ad*xp[k];
  }
}

__global__ __launch_bounds__(TPB) void k_nt_dot(long long N, const double *a, const double *b, double *part)
{
  double s = 0.0, z1 = 0.0, z2 = 0.0;
  for (long long i = (long long)blockIdx.x*TPB + threadIdx.x; i < N; i += (long long)gridDim.x*TPB) s += a[i]*b[i];
  block_sum3(s, z1, z2);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_nt_axpy(long long N, double alpha, const double *x, double *y)
{
  const long long i = (long long)blockIdx.x*blockDim.x + threadIdx.x;
  if (i < N) y[i] += alpha*x[i];
}

__global__ void k_nt_scale(long long N, double t, double *x)
{
  const long long i = (long long)blockIdx.x*blockDim.x + threadIdx.x;
  if (i < N) x[i] *= t;
}

__global__ void k_nt_precond(int n, const double *P, const double *b, double *x)
{
  const int p = blockIdx.x*blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double q[3] = {b[3*(size_t)p], b[3*(size_t)p + 1], b[3*(size_t)p + 2]};
#pragma unroll
  for (int r = 0; r < 3; r ++)
    x[3*(size_t)p + r] = P[(size_t)r*n + p]*q[0] + P[(size_t)(r + 3)*n + p]*q[1] + P[(size_t)(r + 6)*n + p]*q[2];
}

// R = proj(R + dr) (nts.c:1139-1154)
__global__ void k_nt_update(int n, double *R, const double *dr, const double *fri, const double *coh)
{
  const int p = blockIdx.x*blockDim.x + threadIdx.x;
  if (p >= n) return;
  double S[3], Rp[3];
#pragma unroll
  for (int k = 0; k < 3; k ++) S[k] = R[4*(size_t)p + k] + dr[3*(size_t)p + k];
  scf_project(fri[p], coh[p], S, Rp);
#pragma unroll
  for (int k = 0; k < 3; k ++) R[4*(size_t)p + k] = Rp[k];
}

// the Arnoldi scalars of one GMRES cycle live on the device: the Hessenberg
// matrix hh (row-major, (k_dim + 1) x k_dim), the right-hand side rs and the
// Givens rotations c, s, and the last residual estimate
struct NtScal {
  double hh[(NT_KDIM + 1)*NT_KDIM];
  double rs[NT_KDIM + 1], c[NT_KDIM], s[NT_KDIM];
  double rnorm;
};

// the sum of a dot product's block partials, in a fixed order; sqrt of it when root
__global__ __launch_bounds__(TPB) void k_nt_dotfin(const double *part, int np, double *out, int root)
{
  double v = threadIdx.x < np ? part[threadIdx.x] : 0.0, z1 = 0.0, z2 = 0.0;
  block_sum3(v, z1, z2);
  if (threadIdx.x == 0) *out = root ? sqrt(v) : v;
}

// y += sign * (*alpha) x
__global__ void k_nt_axpy_dev(long long N, const double *alpha, double sign, const double *x, double *y)
{
  const long long i = (long long)blockIdx.x*blockDim.x + threadIdx.x;
  if (i < N) y[i] += (sign*(*alpha))*x[i];
}

// x *= 1 / (*t) unless *t is 0
__global__ void k_nt_scale_inv(long long N, const double *t, double *x)
{
  const long long i = (long long)blockIdx.x*blockDim.x + threadIdx.x;
  const double tv = *t;
  if (i < N && tv != 0.0) x[i] *= 1.0/tv;
}

// column i - 1 of the Arnoldi step: the previous rotations, a new one, the
// updated right-hand side and the residual estimate |rs[i]| (flexgmres.c:531-557)
__global__ void k_nt_givens(NtScal *sc, int i)
{
#pragma clang fp contract(off)
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double *hh = sc->hh, *c = sc->c, *s = sc->s, *rs = sc->rs, t;
#define HH(a, b) hh[(a)*NT_KDIM + (b)]
  for (int j = 1; j < i; j ++)
  {
    t = HH(j - 1, i - 1);
    HH(j - 1, i - 1) = s[j - 1]*HH(j, i - 1) + c[j - 1]*t;
    HH(j, i - 1) = -s[j - 1]*t + c[j - 1]*HH(j, i - 1);
  }
  t = HH(i, i - 1)*HH(i, i - 1);
  t += HH(i - 1, i - 1)*HH(i - 1, i - 1);
  double gamma = sqrt(t);
  if (gamma == 0.0) gamma = 1.e-16;
  c[i - 1] = HH(i - 1, i - 1)/gamma;
  s[i - 1] = HH(i, i - 1)/gamma;
  rs[i] = -HH(i, i - 1)*rs[i - 1];
  rs[i] /= gamma;
  rs[i - 1] = c[i - 1]*rs[i - 1];
  HH(i - 1, i - 1) = s[i - 1]*HH(i, i - 1) + c[i - 1]*HH(i - 1, i - 1);
  sc->rnorm = fabs(rs[i]);
#undef HH
}

// the device vectors of one Newton solve
struct NtWork {
  sfgs_ctx *ctx;
  int n;
  long long N;
  double *X, *Y, *P, *b, *dr, *R0, *kry, *pre, *w, *r, *part;
  NtScal *sc;
  NtMat mat;
  int bodyspace = 0;
  int matvecs = 0;
  std::vector<double> hpart;
  double rs0 = 0.0, rnorm_h = 0.0;   // host ends of the small copies

  double *kv(int k) const { return kry + (size_t)k*N; }
  double *pv(int k) const { return pre + (size_t)k*N; }

  int matvec(double alpha, const double *x, double beta, double *y)
  {
    if (bodyspace)
    {
      CHK(bodyspace_product(ctx, x, 3, 0, X, Y, alpha, beta, mat.delta, y));
      matvecs ++;
      return 0;
    }
    const int ns = ctx->nslices, grid = (ns + TPB/SLICE - 1) / (TPB/SLICE);
    if (grid > 0)
      NT_DISPATCH_G(ctx->G_built, auto kern = k_nt_matvec<GG>; hipLaunchKernelGGL(kern, dim3(grid), dim3(TPB), 0,
                       ctx->stream, mat, ns, alpha, x, beta, y));
    matvecs ++;
    HIPCHK(hipGetLastError());
    return 0;
  }

  int dot(const double *a, const double *b2, double &out)
  {
    const int grid = std::max(1, std::min(NT_DOT_BLOCKS, nblk(N)));
    hipLaunchKernelGGL(k_nt_dot, dim3(grid), dim3(TPB), 0, ctx->stream, N, a, b2, part);
    HIPCHK(hipMemcpyAsync(hpart.data(), part, sizeof(double)*(size_t)grid, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    double s = 0.0;
    for (int k = 0; k < grid; k ++) s += hpart[k];
    out = s;
    return 0;
  }

  // x'y (or its square root) into device memory: no read-back
  void dot_dev(const double *a, const double *b2, double *out, int root)
  {
    const int grid = std::max(1, std::min(NT_DOT_BLOCKS, nblk(N)));
    hipLaunchKernelGGL(k_nt_dot, dim3(grid), dim3(TPB), 0, ctx->stream, N, a, b2, part);
    hipLaunchKernelGGL(k_nt_dotfin, dim3(1), dim3(TPB), 0, ctx->stream, part, grid, out, root);
  }

  void axpy_dev(const double *alpha, double sign, const double *x, double *y)
  {
    if (N > 0) hipLaunchKernelGGL(k_nt_axpy_dev, dim3(nblk(N)), dim3(TPB), 0, ctx->stream, N, alpha, sign, x, y);
  }

  void scale_inv_dev(const double *t, double *x)
  {
    if (N > 0) hipLaunchKernelGGL(k_nt_scale_inv, dim3(nblk(N)), dim3(TPB), 0, ctx->stream, N, t, x);
  }

  void axpy(double alpha, const double *x, double *y)
  {
    if (N > 0) hipLaunchKernelGGL(k_nt_axpy, dim3(nblk(N)), dim3(TPB), 0, ctx->stream, N, alpha, x, y);
  }

  void scale(double t, double *x)
  {
    if (N > 0) hipLaunchKernelGGL(k_nt_scale, dim3(nblk(N)), dim3(TPB), 0, ctx->stream, N, t, x);
  }

  int copy(const double *x, double *y)
  {
    if (N > 0) HIPCHK(hipMemcpyAsync(y, x, sizeof(double)*(size_t)N, hipMemcpyDeviceToDevice, ctx->stream));
    return 0;
  }

  void precond(const double *bb, double *x)
  {
    if (n > 0) hipLaunchKernelGGL(k_nt_precond, dim3(nblk(n)), dim3(TPB), 0, ctx->stream, n, P, bb, x);
  }

  // hypre_FlexGMRESSolve (flexgmres.c:260-700) with rel_change off, cf_tol 0,
  // r_tol 0, k_dim 20: x updated in place.  The Arnoldi step runs on the
  // device (dot products reduced there, the Givens update in one thread); the
  // host reads the residual estimate once per iteration and the small
  // Hessenberg system once per cycle.
  int gmres(const double *bb, double *x, int max_iter, int min_iter, double a_tol)
  {
    const int k_dim = NT_KDIM;
    NtScal h;
    double r_norm, b_norm, t;
    int iter = 0, i = 0;
    double *hh = sc->hh;
#define HHD(a, b) (hh + (a)*NT_KDIM + (b))
#define HHH(a, b) h.hh[(a)*NT_KDIM + (b)]
    CHK(copy(bb, kv(0)));
    CHK(matvec(-1.0, x, 1.0, kv(0)));
    CHK(dot(bb, bb, t)); b_norm = std::sqrt(t);
    if (b_norm != 0.0 && !std::isfinite(b_norm)) return 0;   // non-finite input: x untouched
    CHK(dot(kv(0), kv(0), t)); r_norm = std::sqrt(t);
    if (r_norm != 0.0 && !std::isfinite(r_norm)) return 0;
    const double epsilon = a_tol;
    while (iter < max_iter)
    {
      if (r_norm == 0.0) break;
      if (r_norm <= epsilon && iter >= min_iter)
      {
        CHK(copy(bb, r));
        CHK(matvec(-1.0, x, 1.0, r));
        CHK(dot(r, r, t)); r_norm = std::sqrt(t);
        if (r_norm <= epsilon) break;
      }
      rs0 = r_norm;   // rs[0] of the cycle
      HIPCHK(hipMemcpyAsync(sc->rs, &rs0, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
      scale(1.0/r_norm, kv(0));
      i = 0;
      while (i < k_dim && iter < max_iter)
      {
        i ++;
        iter ++;
        HIPCHK(hipMemsetAsync(pv(i - 1), 0, sizeof(double)*(size_t)N, ctx->stream));
        precond(kv(i - 1), pv(i - 1));
        CHK(matvec(1.0, pv(i - 1), 0.0, kv(i)));
        for (int j = 0; j < i; j ++)   // modified Gram-Schmidt
        {
          dot_dev(kv(j), kv(i), HHD(j, i - 1), 0);
          axpy_dev(HHD(j, i - 1), -1.0, kv(j), kv(i));
        }
        dot_dev(kv(i), kv(i), HHD(i, i - 1), 1);
        scale_inv_dev(HHD(i, i - 1), kv(i));
        hipLaunchKernelGGL(k_nt_givens, dim3(1), dim3(64), 0, ctx->stream, sc, i);
        HIPCHK(hipMemcpyAsync(&rnorm_h, &sc->rnorm, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        r_norm = rnorm_h;
        if (r_norm <= epsilon && iter >= min_iter) break;
      }
      HIPCHK(hipMemcpyAsync(&h, sc, sizeof(NtScal), hipMemcpyDeviceToHost, ctx->stream));
      HIPCHK(hipStreamSynchronize(ctx->stream));
      HIPCHK(hipGetLastError());
      h.rs[i - 1] = h.rs[i - 1]/HHH(i - 1, i - 1);
      for (int k = i - 2; k >= 0; k --)
      {
        t = 0.0;
        for (int j = k + 1; j < i; j ++) t -= HHH(k, j)*h.rs[j];
        t += h.rs[k];
        h.rs[k] = t/HHH(k, k);
      }
      CHK(copy(pv(i - 1), w));
      scale(h.rs[i - 1], w);
      for (int j = i - 2; j >= 0; j --) axpy(h.rs[j], pv(j), w);
      axpy(1.0, w, x);
      if (r_norm <= epsilon && iter >= min_iter)
      {
        CHK(copy(bb, r));
        CHK(matvec(-1.0, x, 1.0, r));
        CHK(dot(r, r, t)); r_norm = std::sqrt(t);
        if (r_norm <= epsilon) break;
        CHK(copy(r, kv(0)));   // false convergence
        i = 0;
      }
      for (int j = i; j > 0; j --)
      {
        h.rs[j - 1] = -h.s[j - 1]*h.rs[j];
        h.rs[j] = h.c[j - 1]*h.rs[j];
      }
      if (i) axpy(h.rs[i] - 1.0, kv(i), kv(i));
      for (int j = i - 1; j > 0; j --) axpy(h.rs[j], kv(j), kv(i));
      if (i)
      {
        axpy(h.rs[0] - 1.0, kv(0), kv(0));
        axpy(1.0, kv(i), kv(0));
      }
    }
#undef HHD
#undef HHH
    HIPCHK(hipGetLastError());
    return 0;
  }
};

} // namespace

void sfgs_newton_release(sfgs_ctx *ctx)
{
  release(ctx->nt_mem);
}

extern "C" {

// NEWTON_Solve (nts.c:1333-1379) with LOCDYN_ON
int SFGS_Newton_Solve(sfgs_ctx *ctx, const SFGS_NEWTON_PARAM *par, double *merhist, int *mvhist,
                      SFGS_NEWTON_REPORT *rep)
{
  if (!ctx || !par) return SFGS_E_ARG;
  if (!(par->meritval > 0.0) || par->maxiter <= 0 || par->linmaxiter <= 0 || par->maxmatvec <= 0 ||
      !(par->epsilon > 0.0) || !(par->delta >= 0.0) || !(par->theta > 0.0 && par->theta <= 1.0) ||
      !(par->omega > 0.0) || (par->linver != SFGS_NT_GMRES && par->linver != SFGS_NT_DIAG) ||
      par->reldelta < 0 || par->reldelta > 3 || (par->bodyspace != 0 && par->bodyspace != 1)) return SFGS_E_ARG;
  if (!ctx->assembled) { ctx->err = "solve before LOCDYN assembly"; return SFGS_E_STATE; }
  if (ctx->diag_only) { ctx->err = "the store holds a penalty solve, not an assembled system"; return SFGS_E_STATE; }
  if (ctx->nscenes > 1) { ctx->err = "the system holds independent scenes"; return SFGS_E_STATE; }
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int n = ctx->nc;
  const long long N = 3LL*n;
  const size_t N1 = (size_t)std::max(N, 1LL), n1 = (size_t)std::max(n, 1);

  // one allocation: X, Y, P (9 per row), b, dr, w, r (3 per row), R0 (4 per
  // row), the Krylov and the preconditioned vectors, dot partials
  const size_t scal_words = (sizeof(NtScal) + 7)/8;
  const size_t words = 27*n1 + 4*N1 + 4*n1 + (size_t)(2*NT_KDIM + 1)*N1 + NT_DOT_BLOCKS + scal_words;
  CHK(ensure(ctx, ctx->nt_mem, words));
  NtWork wk;
  double *m0 = ctx->nt_mem.p;
  wk.ctx = ctx; wk.n = n; wk.N = N;
  wk.X = m0; wk.Y = wk.X + 9*n1; wk.P = wk.Y + 9*n1; wk.b = wk.P + 9*n1; wk.dr = wk.b + N1; wk.w = wk.dr + N1;
  wk.r = wk.w + N1; wk.R0 = wk.r + N1; wk.kry = wk.R0 + 4*n1; wk.pre = wk.kry + (size_t)(NT_KDIM + 1)*N1;
  wk.part = wk.pre + (size_t)NT_KDIM*N1;
  wk.sc = (NtScal*)(wk.part + NT_DOT_BLOCKS);
  wk.hpart.resize(NT_DOT_BLOCKS);
  HIPCHK(hipMemsetAsync(m0, 0, sizeof(double)*words, st));
  if (n > 0) HIPCHK(hipMemcpyAsync(wk.R0, ctx->R.p, sizeof(double)*4*(size_t)n, hipMemcpyDeviceToDevice, st));
  CHK(ensure(ctx, ctx->failcount, 4));
  HIPCHK(hipMemsetAsync(ctx->failcount.p, 0, sizeof(int)*4, st));

  // W norm of the relative delta (LOCDYN_avgWii / minWii / maxWii, ldy.c:1146-1193)
  double W_norm = 1.0;
  if (par->reldelta != 0 && n > 0)
  {
    std::vector<double> d(3*(size_t)n);
    for (int e = 0; e < 3; e ++)
      HIPCHK(hipMemcpyAsync(d.data() + (size_t)e*n, ctx->Wd.p + (size_t)(4*e)*n, sizeof(double)*(size_t)n,
                            hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (par->reldelta == 1)
    {
      double sum = 0.0, div = 0.0;
      for (int p = 0; p < n; p ++) { sum += d[p] + d[(size_t)n + p] + d[2*(size_t)n + p]; div += 3.0; }
      W_norm = sum/div;
    }
    else if (par->reldelta == 2) W_norm = *std::min_element(d.begin(), d.end());
    else W_norm = std::max(0.0, *std::max_element(d.begin(), d.end()));
  }
  else if (par->reldelta == 2) W_norm = DBL_MAX;
  else if (par->reldelta == 3) W_norm = 0.0;
  const double delta = par->delta*W_norm;

  wk.mat = NtMat{ctx->slice_row0.p, ctx->slice_nrows.p, ctx->rowlen.p, ctx->slice_off.p, ctx->W.p, ctx->wcol.p,
                 ctx->Wd.p, wk.X, wk.Y, n, delta};
  NtRows rows{n, ctx->R.p, ctx->U.p, ctx->Wd.p, ctx->pfri.p, ctx->pcoh.p, ctx->shift.p, wk.X, wk.Y, wk.P, wk.b, wk.dr,
              ctx->failcount.p};

  wk.bodyspace = par->bodyspace;
  if (par->bodyspace && !ctx->from_bodies)
  {
    ctx->err = "locdyn OFF needs a system assembled from bodies";
    return SFGS_E_STATE;
  }
  if (!par->bodyspace && ctx->upmin)
  {
    ctx->err = "locdyn ON needs the off-diagonal blocks (SFGS_Locdyn_Update_Begin)";
    return SFGS_E_STATE;
  }
  // U = W R + B (U_WR_B, nts.c:255-431), then the merit (nts.c:929, 1199); in
  // body space W R is formed through the bodies
  auto merit_of = [&](double &m) -> int {
    if (!par->bodyspace) return SFGS_Merit(ctx, 1, &m, nullptr);
    CHK(bodyspace_product(ctx, ctx->R.p, 4, 1, nullptr, nullptr, 0.0, 0.0, 0.0, nullptr));
    return SFGS_Merit(ctx, 0, &m, nullptr);
  };
  double merit = 0.0;
  CHK(merit_of(merit));
  int iters = 0, nan = 0;
  const bool diag = par->linver == SFGS_NT_DIAG;
  while (iters < par->maxiter && (diag || wk.matvecs < par->maxmatvec) && merit > par->meritval)
  {
    if (n > 0)
      hipLaunchKernelGGL(k_nt_lin, dim3(nblk(n)), dim3(TPB), 0, st, rows, diag ? 1 : 0, par->omega, delta, par->theta,
                         ctx->R.p);
    HIPCHK(hipGetLastError());
    if (!diag)
    {
      double bb = 0.0;
      CHK(wk.dot(wk.b, wk.b, bb));
      CHK(wk.gmres(wk.b, wk.dr, par->linmaxiter, 1, par->epsilon*std::sqrt(bb)));
      if (n > 0)
        hipLaunchKernelGGL(k_nt_update, dim3(nblk(n)), dim3(TPB), 0, st, n, ctx->R.p, wk.dr, ctx->pfri.p, ctx->pcoh.p);
      HIPCHK(hipGetLastError());
    }
    CHK(merit_of(merit));
    if (std::isnan(merit)) { nan = 1; break; }
    if (merhist) merhist[iters] = merit;
    if (mvhist) mvhist[iters] = diag ? 0 : wk.matvecs;
    iters ++;
  }
  int singular = 0;
  HIPCHK(hipMemcpyAsync(&singular, ctx->failcount.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  SFGS_NEWTON_REPORT r;
  std::memset(&r, 0, sizeof(r));
  r.iters = iters;
  r.matvecs = wk.matvecs;
  r.converged = !nan && merit <= par->meritval;
  r.singular = singular;
  r.merit = merit;
  r.W_norm = W_norm;
  if (!r.converged && par->gsflag && par->bodyspace)
  {
    // the reference aborts here (nts.c:1370-1374: no Gauss-Seidel without the
    // assembled blocks); this library reports the failure instead
    if (rep) *rep = r;
    ctx->err = "NEWTON_SOLVER failed in body space mode";
    return SFGS_E_STATE;
  }
  if (!r.converged && par->gsflag)
  {
    // R back to its initial value; GAUSS_SEIDEL_Create (1.0, maxiter, meritval,
    // GS_FAILURE_CONTINUE, 1E-9, 100, DS_SEMISMOOTH_NEWTON) (nts.c:1363)
    if (n > 0) HIPCHK(hipMemcpyAsync(ctx->R.p, wk.R0, sizeof(double)*4*(size_t)n, hipMemcpyDeviceToDevice, st));
    SFGS_PARAM gp;
    std::memset(&gp, 0, sizeof(gp));
    gp.epsilon = 1.0; gp.maxiter = par->maxiter; gp.meritval = par->meritval; gp.failure = SFGS_FAILURE_CONTINUE;
    gp.diagepsilon = 1e-9; gp.diagmaxiter = 100; gp.diagsolver = SFGS_DS_SEMISMOOTH_NEWTON;
    gp.precision = SFGS_PREC_FP64;
    SFGS_REPORT gr;
    CHK(SFGS_Solve(ctx, &gp, nullptr, nullptr, &gr));
    r.gs_fallback = 1;
    r.gs_iters = gr.iters;
    r.merit = gr.merit;
  }
  if (rep) *rep = r;
  return 0;
}

} // extern "C"
