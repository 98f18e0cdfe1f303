// sfgs_kernels.hip -- LOCDYN assembly, contact-graph colouring, colour-major
// block Gauss-Seidel sweeps, merit and body update on MI355X (gfx950).
//
// Data layout (all device resident, rebuilt every step):
//  * natural order (contact index i): the uploaded contact records and R, U;
//  * colour-major order (row p = iperm[i]): rows sorted by (colour, -row length,
//    i); per-row scalars in SoA planes (Wd[e*n + p], B[k*n + p], U[k*n + p]),
//    reactions R as 4 doubles per row (32-byte aligned gathers);
//  * off-diagonal blocks in SELL-64: a slice is 64 consecutive rows of one
//    colour (one wavefront, one row per lane); slot k of lane l stores element e
//    of its 3x3 block at W[9*off + (9k + e)*64 + l] and the neighbour's row at
//    wcol[off + 64k + l], so each load instruction of a wave reads 512
//    contiguous bytes.
// One kernel launch sweeps one colour: rows of a colour are not coupled, so the
// launch order of colours reproduces a serial Gauss-Seidel sweep over the
// colour-major row order exactly.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include <cfloat>
#include <climits>
#include <algorithm>
#include <mutex>
#include <string>
#include "sfgs_ctx.hpp"
#include "sfgs_math.hpp"
#include "sfgs_prims.hpp"
#include "sfgs_sweep.hpp"
#include "sfgs_host.hpp"

using namespace sfgs;

// =====================================================================
// device helpers
// =====================================================================

__device__ __forceinline__ unsigned hash32(unsigned x)
{
  // 32-bit integer mixer "lowbias32" (C. Wellons, hash-prospector)
  x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
  return x;
}

__device__ __forceinline__ bool couples(const int *kind, int b) { return b >= 0 && kind[b] == 0; }

// H operator of a rigid body at referential point X (bod.c:216-228):
// H = [ -base' R skew(X - X0) | base' ], returned as Hr (3x3) and the frame.
__device__ __forceinline__ void operator_Hr(const double *R, const double *X0, const double *X,
                                            const double *base, double *Hr)
{
  double a0 = X[0] - X0[0], a1 = X[1] - X0[1], a2 = X[2] - X0[2];
  double Q[9];   // Q = base' R
#pragma unroll
  for (int c = 0; c < 3; c ++)
#pragma unroll
    for (int r = 0; r < 3; r ++)
      Q[r + 3*c] = base[3*r]*R[3*c] + base[3*r + 1]*R[3*c + 1] + base[3*r + 2]*R[3*c + 2];
  // Hr = -Q skew(a): columns  -(a2 Q1 - a1 Q2), -(a0 Q2 - a2 Q0), -(a1 Q0 - a0 Q1)
#pragma unroll
  for (int r = 0; r < 3; r ++)
  {
    double q0 = Q[r], q1 = Q[r + 3], q2 = Q[r + 6];
    Hr[r]     = a1*q2 - a2*q1;
    Hr[r + 3] = a2*q0 - a0*q2;
    Hr[r + 6] = a0*q1 - a1*q0;
  }
}

// P = H1 iM H2' = Hr1 J Hr2' + minv base1' base2  (3x3, column-major)
__device__ __forceinline__ void h_im_ht(const double *Hr1, const double *b1, const double *Hr2, const double *b2,
                                        const double *J, double minv, double *P)
{
  double G[9];   // G = J Hr2'
#pragma unroll
  for (int c = 0; c < 3; c ++)
#pragma unroll
    for (int r = 0; r < 3; r ++)
      G[r + 3*c] = J[r]*Hr2[c] + J[r + 3]*Hr2[c + 3] + J[r + 6]*Hr2[c + 6];
#pragma unroll
  for (int c = 0; c < 3; c ++)
#pragma unroll
    for (int r = 0; r < 3; r ++)
      P[r + 3*c] = (Hr1[r]*G[3*c] + Hr1[r + 3]*G[3*c + 1] + Hr1[r + 6]*G[3*c + 2])
                 + minv*(b1[3*r]*b2[3*c] + b1[3*r + 1]*b2[3*c + 1] + b1[3*r + 2]*b2[3*c + 2]);
}

// y = H u, u = [W v]
__device__ __forceinline__ void h_u(const double *Hr, const double *base, const double *u, double *y)
{
#pragma unroll
  for (int r = 0; r < 3; r ++)
    y[r] = Hr[r]*u[0] + Hr[r + 3]*u[1] + Hr[r + 6]*u[2]
         + base[3*r]*u[3] + base[3*r + 1]*u[4] + base[3*r + 2]*u[5];
}

// cyclic Jacobi eigenvalues of a symmetric 3x3 (upper triangle), maximum returned
__device__ double eig3_max(const double *W)
{
  double a00 = W[0], a01 = W[3], a02 = W[6], a11 = W[4], a12 = W[7], a22 = W[8];
  double a[3][3] = {{a00, a01, a02}, {a01, a11, a12}, {a02, a12, a22}};
  for (int sweep = 0; sweep < 64; sweep ++)
  {
    double off = a[0][1]*a[0][1] + a[0][2]*a[0][2] + a[1][2]*a[1][2];
    double dia = a[0][0]*a[0][0] + a[1][1]*a[1][1] + a[2][2]*a[2][2];
    if (off <= 1e-34*dia || off == 0.0) break;
    for (int p = 0; p < 2; p ++)
      for (int q = p + 1; q < 3; q ++)
      {
        if (a[p][q] == 0.0) continue;
        double theta = (a[q][q] - a[p][p]) / (2.0*a[p][q]);
        double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta*theta + 1.0));
        double c = 1.0/sqrt(t*t + 1.0), s = t*c;
        for (int k = 0; k < 3; k ++)
        {
          double akp = a[k][p], akq = a[k][q];
          a[k][p] = c*akp - s*akq;
          a[k][q] = s*akp + c*akq;
        }
        for (int k = 0; k < 3; k ++)
        {
          double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c*apk - s*aqk;
          a[q][k] = s*apk + c*aqk;
        }
      }
  }
  return fmax(a[0][0], fmax(a[1][1], a[2][2]));
}

// deterministic block sum of two values; thread 0 returns them
__device__ __forceinline__ void block_sum2(double &a, double &b)
{
  __shared__ double sa[TPB/64], sb[TPB/64];
  a = wave_sum(a); b = wave_sum(b);
  int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { sa[w] = a; sb[w] = b; }
  __syncthreads();
  if (threadIdx.x == 0)
  {
    a = sa[0]; b = sb[0];
    for (int k = 1; k < TPB/64; k ++) { a += sa[k]; b += sb[k]; }
  }
}

// =====================================================================
// incidence and adjacency (LOCDYN_Insert, ldy.c:492-568)
// =====================================================================

__global__ void k_inc_keys(int nc, const int *master, const int *slave, const int *kind, int *keys, int *vals)
{
  int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i >= nc) return;
  int m = master[i], s = slave[i];
  keys[2*i] = couples(kind, m) ? m : INT_MAX;
  keys[2*i + 1] = (couples(kind, s) && s != m) ? s : INT_MAX;
  vals[2*i] = i;
  vals[2*i + 1] = i;
}

__global__ void k_inc_count(int n2, const int *keys, int *count)
{
  int q = blockIdx.x*blockDim.x + threadIdx.x;
  if (q >= n2) return;
  int b = keys[q];
  if (b != INT_MAX) atomicAdd(&count[b], 1);
}

__global__ void k_rowlen(int nc, const int *master, const int *slave, const int *kind, const int *bptr, int *len)
{
  int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i >= nc) return;
  int m = master[i], s = slave[i], k = 0;
  if (couples(kind, m)) k += bptr[m + 1] - bptr[m] - 1;
  if (couples(kind, s) && s != m) k += bptr[s + 1] - bptr[s] - 1;
  len[i] = k;
}

__global__ void k_adj_fill(int nc, const int *master, const int *slave, const int *kind, const int *bptr,
                           const int *blist, const int *rowptr, int *col, int *via)
{
  int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i >= nc) return;
  int p = rowptr[i], side[2] = {master[i], slave[i]};
  for (int t = 0; t < 2; t ++)
  {
    int b = side[t];
    if (!couples(kind, b) || (t == 1 && b == side[0])) continue;
    for (int q = bptr[b]; q < bptr[b + 1]; q ++)
    {
      int j = blist[q];
      if (j != i) { col[p] = j; via[p] = b; p ++; }
    }
  }
}

// =====================================================================
// Jones-Plassmann colouring: priority (row length << 32 | hash(v)), ties by
// larger index; a vertex is coloured when no uncoloured neighbour outranks it
// and takes the smallest colour absent from its coloured neighbours.  Reads
// cin only (the previous round), writes cout: the result equals greedy
// first-fit in descending priority and does not depend on scheduling.
// =====================================================================

#define SFGS_MAX_COLOURS 256

__global__ void k_jp_round(int n, const int *rowptr, const int *col, const int *cin, int *cout,
                           int *remaining, int *maxcol)
{
  int v = blockIdx.x*blockDim.x + threadIdx.x;
  if (v >= n || cin[v] >= 0) return;
  unsigned long long pv = ((unsigned long long)(unsigned)(rowptr[v + 1] - rowptr[v]) << 32) | hash32((unsigned)v);
  unsigned long long mask[SFGS_MAX_COLOURS/64] = {0ull, 0ull, 0ull, 0ull};
  for (int p = rowptr[v]; p < rowptr[v + 1]; p ++)
  {
    int u = col[p], cu = cin[u];
    if (cu < 0)
    {
      unsigned long long pu = ((unsigned long long)(unsigned)(rowptr[u + 1] - rowptr[u]) << 32) | hash32((unsigned)u);
      if (pu > pv || (pu == pv && u > v)) { atomicAdd(remaining, 1); return; }
    }
    else if (cu < SFGS_MAX_COLOURS) mask[cu >> 6] |= 1ull << (cu & 63);
  }
  int c = SFGS_MAX_COLOURS;
  for (int w = 0; w < SFGS_MAX_COLOURS/64; w ++)
    if (~mask[w]) { c = 64*w + __ffsll((long long)~mask[w]) - 1; break; }
  cout[v] = c;
  atomicMax(maxcol, c);
}

__global__ void k_copy_int(int n, const int *a, int *b)
{
  int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

__global__ void k_fill_int(int n, int v, int *b)
{
  int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i < n) b[i] = v;
}

// monotone map of a double onto unsigned 64-bit integers (a < b <=> key(a) < key(b))
__device__ __forceinline__ unsigned long long order_key(double d)
{
  const unsigned long long b = (unsigned long long)__double_as_longlong(d);
  return (b >> 63) ? ~b : b | 0x8000000000000000ULL;
}

__device__ __forceinline__ double order_val(unsigned long long k)
{
  return __longlong_as_double((long long)((k >> 63) ? k & 0x7fffffffffffffffULL : ~k));
}

__global__ void k_bbox_init(unsigned long long *bb)
{
  if (threadIdx.x < 6) bb[threadIdx.x] = threadIdx.x < 3 ? ~0ULL : 0ULL;
}

// bounding box of the positions of the bodies that take part in the dynamics
// (kind 0), exact and order-independent: bb[0..2] minima, bb[3..5] maxima of
// order_key (k_bbox_init first)
__global__ __launch_bounds__(TPB) void k_bbox(int nb, const int *kind, const double *x, unsigned long long *bb)
{
  int b = blockIdx.x*blockDim.x + threadIdx.x;
  unsigned long long lo[3] = {~0ULL, ~0ULL, ~0ULL}, hi[3] = {0ULL, 0ULL, 0ULL};
  if (b < nb && kind[b] == 0)
  {
#pragma unroll
    for (int k = 0; k < 3; k ++) lo[k] = hi[k] = order_key(x[3*(size_t)b + k]);
  }
  // wavefront, then workgroup minima and maxima; one atomic per value and workgroup
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
  {
#pragma unroll
    for (int k = 0; k < 3; k ++)
    {
      lo[k] = min(lo[k], (unsigned long long)__shfl_xor(lo[k], o, 64));
      hi[k] = max(hi[k], (unsigned long long)__shfl_xor(hi[k], o, 64));
    }
  }
  __shared__ unsigned long long sh[TPB/64][6];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
  {
#pragma unroll
    for (int k = 0; k < 3; k ++) { sh[w][k] = lo[k]; sh[w][3 + k] = hi[k]; }
  }
  __syncthreads();
  if (threadIdx.x < 6)
  {
    const int k = threadIdx.x;
    unsigned long long v = sh[0][k];
    for (int q = 1; q < TPB/64; q ++) v = k < 3 ? min(v, sh[q][k]) : max(v, sh[q][k]);
    if (k < 3) { if (v != ~0ULL) atomicMin(bb + k, v); }
    else if (v != 0ULL) atomicMax(bb + k, v);
  }
}

// the 10 low bits of v spread to every third bit (Morton code of one axis)
__device__ __forceinline__ unsigned spread3(unsigned v)
{
  unsigned x = v & 0x3ffu;
  x = (x | (x << 16)) & 0x030000ffu;
  x = (x | (x << 8)) & 0x0300f00fu;
  x = (x | (x << 4)) & 0x030c30c3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}

// sort key of row i, by order mode: 0 (scene, colour, -row length); 1 (scene,
// colour), the stable sort keeping contact order inside a colour; 2 (scene,
// colour, master body / 64, -row length), which keeps rows of nearby bodies
// together while slices stay of even length; 3 (scene, colour, Z-order of the
// master body's position on a 1024^3 grid over the bounding box bb of the
// bodies), which keeps rows of bodies that are close in space together.
// scene = 0 without scenes
__global__ void k_sort_keys(int n, const int *colour, const int *rowptr, const int *scene, int mode,
                            const int *master, const double *x, const unsigned long long *bb,
                            unsigned long long *keys, int *vals)
{
  int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int len = rowptr[i + 1] - rowptr[i];
  const unsigned long long sc = scene ? (unsigned long long)(unsigned)scene[i] : 0ULL;
  const unsigned long long col = (unsigned long long)(unsigned)colour[i];
  if (mode == 3)
  {
    const double *p = x + 3*(size_t)master[i];
    unsigned q[3];
#pragma unroll
    for (int k = 0; k < 3; k ++)
    {
      const double lo = order_val(bb[k]), hi = order_val(bb[3 + k]);
      const double t = hi > lo ? (p[k] - lo) / (hi - lo) * 1024.0 : 0.0;
      q[k] = (unsigned)fmin(fmax(t, 0.0), 1023.0);
    }
    const unsigned long long z = (unsigned long long)(spread3(q[0]) | (spread3(q[1]) << 1) | (spread3(q[2]) << 2));
    keys[i] = (sc << 40) | (col << 32) | z;
  }
  else if (mode == 2)
  {
    const unsigned long long blk = (unsigned long long)((unsigned)master[i] >> 6) & 0xffffffULL;
    keys[i] = (sc << 40) | (col << 32) | (blk << 8) | (0xffu - (unsigned)min(len, 0xff));
  }
  else
  {
    const unsigned l24 = mode == 1 ? 0u : (unsigned)min(len, 0xffffff);
    keys[i] = (sc << 32) | (col << 24) | (0xffffffu - (mode == 1 ? 0xffffffu : l24));
  }
  vals[i] = i;
}

// scene of each contact (its master body's); a coupling slave of another scene is an error
__global__ void k_row_scene(int n, const int *master, const int *slave, const int *kind, const int *body_scene,
                            int *scene, int *bad)
{
  int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i >= n) return;
  int sc = body_scene[master[i]], s = slave[i];
  scene[i] = sc;
  if (s >= 0 && kind[s] == 0 && body_scene[s] != sc) atomicAdd(bad, 1);
}

// rows per (scene, colour) group
__global__ void k_group_count(int n, const int *colour, const int *scene, int ncolours, int *count)
{
  int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i < n) atomicAdd(&count[scene[i]*ncolours + colour[i]], 1);
}

__global__ void k_iperm(int n, const int *perm, int *iperm)
{
  int p = blockIdx.x*blockDim.x + threadIdx.x;
  if (p < n) iperm[perm[p]] = p;
}

// rows per colour: block histogram in LDS, one global add per colour and block
__global__ void k_colour_count(int n, const int *colour, int *count, int ncolours)
{
  __shared__ int h[SFGS_MAX_COLOURS];
  for (int c = threadIdx.x; c < ncolours; c += blockDim.x) h[c] = 0;
  __syncthreads();
  int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i < n) atomicAdd(&h[colour[i]], 1);
  __syncthreads();
  for (int c = threadIdx.x; c < ncolours; c += blockDim.x) if (h[c]) atomicAdd(&count[c], h[c]);
}

__global__ void k_slice_len(int ns, const int *row0, const int *nrows, const int *perm, const int *rowptr, int G,
                            int *slen, long long *slen_ll)
{
  int s = blockIdx.x*blockDim.x + threadIdx.x;
  if (s >= ns) return;
  int len = 0;   // longest row of the slice
  for (int r = 0; r < nrows[s]; r ++)
  {
    int i = perm[row0[s] + r];
    len = max(len, rowptr[i + 1] - rowptr[i]);
  }
  int nslot = (len + G - 1) / G;   // slots of one block per lane
  slen[s] = nslot;
  slen_ll[s] = (long long)nslot * SLICE;
}

// =====================================================================
// LOCDYN assembly (ldy.c:364-397, 626-844)
// =====================================================================

struct BodyView {
  const int *kind;
  const double *rot, *x0, *jinv, *minv, *velo, *velo0;
};

struct ConView {
  const int *master, *slave;
  const double *mpnt, *spnt, *base, *gap, *fri, *rest, *coh, *Rn, *Un;
};

struct RowsOut {
  int n;
  int *rowlen;
  double *Wd, *A, *B, *V, *rho, *pfri, *pcoh, *shift, *R, *U;
  unsigned char *sep;
  double *energy_part;
  int *singular;
};

// per permuted row: V, B (ldy.c:364-397), W_ii, rho, A, free energy (ldy.c:654-736)
__global__ __launch_bounds__(TPB) void k_rows(BodyView bv, ConView cv, RowsOut o, const int *perm,
                                               const int *rowptr, double step, int dynamic)
{
  int p = blockIdx.x*blockDim.x + threadIdx.x;
  double e_part = 0.0, dummy = 0.0;
  if (p < o.n)
  {
    int i = perm[p], n = o.n;
    int m = cv.master[i], s = cv.slave[i];
    const double *base = cv.base + 9*(size_t)i;
    double bs[9];
#pragma unroll
    for (int k = 0; k < 9; k ++) bs[k] = base[k];
    double Hr[9], Wi[9], P[9], y[3], Vv[3], Bv[3];

    operator_Hr(bv.rot + 9*(size_t)m, bv.x0 + 3*(size_t)m, cv.mpnt + 3*(size_t)i, bs, Hr);
    h_u(Hr, bs, bv.velo0 + 6*(size_t)m, Vv);
    h_u(Hr, bs, bv.velo + 6*(size_t)m, Bv);
#pragma unroll
    for (int k = 0; k < 9; k ++) Wi[k] = 0.0;
    if (couples(bv.kind, m))
    {
      h_im_ht(Hr, bs, Hr, bs, bv.jinv + 9*(size_t)m, bv.minv[m], P);
#pragma unroll
      for (int k = 0; k < 9; k ++) Wi[k] += P[k];
    }
    if (s >= 0)
    {
      operator_Hr(bv.rot + 9*(size_t)s, bv.x0 + 3*(size_t)s, cv.spnt + 3*(size_t)i, bs, Hr);
      h_u(Hr, bs, bv.velo0 + 6*(size_t)s, y);
      Vv[0] -= y[0]; Vv[1] -= y[1]; Vv[2] -= y[2];
      h_u(Hr, bs, bv.velo + 6*(size_t)s, y);
      Bv[0] -= y[0]; Bv[1] -= y[1]; Bv[2] -= y[2];
      if (couples(bv.kind, s))
      {
        h_im_ht(Hr, bs, Hr, bs, bv.jinv + 9*(size_t)s, bv.minv[s], P);
#pragma unroll
        for (int k = 0; k < 9; k ++) Wi[k] += P[k];
      }
    }
#pragma unroll
    for (int k = 0; k < 9; k ++) Wi[k] *= step;

    double lmax = eig3_max(Wi), Ai[9];
    if (inverse3(Wi, Ai)) atomicAdd(o.singular, 1);
#pragma unroll
    for (int k = 0; k < 9; k ++) { o.Wd[(size_t)k*n + p] = Wi[k]; o.A[(size_t)k*n + p] = Ai[k]; }
#pragma unroll
    for (int k = 0; k < 3; k ++) { o.B[(size_t)k*n + p] = Bv[k]; o.V[(size_t)k*n + p] = Vv[k]; }
    o.rho[p] = 1.0 / lmax;
    double g = cv.gap[i], e = cv.rest[i];
    o.pfri[p] = cv.fri[i];
    o.pcoh[p] = cv.coh[i];
    o.shift[p] = dynamic ? e * fmin(Vv[2], 0.0) : fmax(g, 0.0) / step;
    o.sep[p] = (dynamic && g > 0.0) ? 1 : 0;
    o.rowlen[p] = rowptr[i + 1] - rowptr[i];
#pragma unroll
    for (int k = 0; k < 3; k ++)
    {
      o.R[4*(size_t)p + k] = cv.Rn[3*(size_t)i + k];
      o.U[(size_t)k*n + p] = cv.Un[3*(size_t)i + k];
    }
    o.R[4*(size_t)p + 3] = 0.0;
    double x0 = Ai[0]*Bv[0] + Ai[3]*Bv[1] + Ai[6]*Bv[2];
    double x1 = Ai[1]*Bv[0] + Ai[4]*Bv[1] + Ai[7]*Bv[2];
    double x2 = Ai[2]*Bv[0] + Ai[5]*Bv[1] + Ai[8]*Bv[2];
    e_part = x0*Bv[0] + x1*Bv[1] + x2*Bv[2];
  }
  block_sum2(e_part, dummy);
  if (threadIdx.x == 0) o.energy_part[blockIdx.x] = e_part;
}

struct SellOut {
  const int *row0, *nrows;
  const long long *off;
  double *W;
  int *wcol;
};

// off-diagonal blocks W_ij = h s_i s_j H_i^b iM_b H_j^b' (ldy.c:738-844).  G
// lanes serve one row: lane g of row r (lane l = rG + g) holds blocks k = mG + g
// in slot m, element e at W[9 off + (9m + e) 64 + l].  The block of a pair is
// formed with the lower contact index as the "upper" side and transposed for
// the other, as the reference's symmetry copy (ldy.c:830-844).
template <int G>
__global__ __launch_bounds__(TPB) void k_offdiag(BodyView bv, ConView cv, SellOut so, const int *perm,
                                                  const int *iperm, const int *rowptr, const int *col,
                                                  const int *via, int nslices, double step)
{
  int s = blockIdx.x*(TPB/SLICE) + (threadIdx.x >> 6), l = threadIdx.x & 63, r = l / G, g = l % G;
  if (s >= nslices || r >= so.nrows[s]) return;
  int p = so.row0[s] + r, i = perm[p];
  long long off = so.off[s];
  double *Wb = so.W + 9*off + l;
  int *cb = so.wcol + off + l;
  for (int k = g, m = 0; k < rowptr[i + 1] - rowptr[i]; k += G, m ++)
  {
    int q = rowptr[i] + k, j = col[q], b = via[q];
    int lo = i < j ? i : j, hi = i < j ? j : i;
    double sl = cv.master[lo] == b ? 1.0 : -1.0, sh = cv.master[hi] == b ? 1.0 : -1.0;
    const double *xl = sl > 0.0 ? cv.mpnt + 3*(size_t)lo : cv.spnt + 3*(size_t)lo;
    const double *xh = sh > 0.0 ? cv.mpnt + 3*(size_t)hi : cv.spnt + 3*(size_t)hi;
    double bl[9], bh[9], Hl[9], Hh[9], P[9];
#pragma unroll
    for (int e = 0; e < 9; e ++) { bl[e] = cv.base[9*(size_t)lo + e]; bh[e] = cv.base[9*(size_t)hi + e]; }
    operator_Hr(bv.rot + 9*(size_t)b, bv.x0 + 3*(size_t)b, xl, bl, Hl);
    operator_Hr(bv.rot + 9*(size_t)b, bv.x0 + 3*(size_t)b, xh, bh, Hh);
    h_im_ht(Hl, bl, Hh, bh, bv.jinv + 9*(size_t)b, bv.minv[b], P);
    double f = step * sl * sh;
    if (i == lo)
    {
#pragma unroll
      for (int e = 0; e < 9; e ++) Wb[(9*(size_t)m + e)*SLICE] = f * P[e];
    }
    else
    {
#pragma unroll
      for (int rr = 0; rr < 3; rr ++)
#pragma unroll
        for (int c = 0; c < 3; c ++)
          Wb[(9*(size_t)m + rr + 3*c)*SLICE] = f * P[c + 3*rr];
    }
    cb[(size_t)m*SLICE] = iperm[j];
  }
}

// ---- installation of a pre-assembled system (natural CSR) ----

struct SysIn {
  const int *rowptr, *col;
  const double *W, *Wd, *A, *B, *V, *rho, *fri, *rest, *coh, *gap, *Rn, *Un;
};

__global__ __launch_bounds__(TPB) void k_rows_sys(SysIn si, RowsOut o, const int *perm, double step, int dynamic)
{
  int p = blockIdx.x*blockDim.x + threadIdx.x;
  if (p >= o.n) return;
  int i = perm[p], n = o.n;
#pragma unroll
  for (int k = 0; k < 9; k ++)
  {
    o.Wd[(size_t)k*n + p] = si.Wd[9*(size_t)i + k];
    o.A[(size_t)k*n + p] = si.A[9*(size_t)i + k];
  }
#pragma unroll
  for (int k = 0; k < 3; k ++)
  {
    o.B[(size_t)k*n + p] = si.B[3*(size_t)i + k];
    o.V[(size_t)k*n + p] = si.V[3*(size_t)i + k];
    o.R[4*(size_t)p + k] = si.Rn[3*(size_t)i + k];
    o.U[(size_t)k*n + p] = si.Un[3*(size_t)i + k];
  }
  o.R[4*(size_t)p + 3] = 0.0;
  double g = si.gap[i], e = si.rest[i], v2 = si.V[3*(size_t)i + 2];
  o.rho[p] = si.rho[i];
  o.pfri[p] = si.fri[i];
  o.pcoh[p] = si.coh[i];
  o.shift[p] = dynamic ? e * fmin(v2, 0.0) : fmax(g, 0.0) / step;
  o.sep[p] = (dynamic && g > 0.0) ? 1 : 0;
  o.rowlen[p] = si.rowptr[i + 1] - si.rowptr[i];
}

template <int G>
__global__ __launch_bounds__(TPB) void k_offdiag_sys(SysIn si, SellOut so, const int *perm, const int *iperm, int nslices)
{
  int s = blockIdx.x*(TPB/SLICE) + (threadIdx.x >> 6), l = threadIdx.x & 63, r = l / G, g = l % G;
  if (s >= nslices || r >= so.nrows[s]) return;
  int p = so.row0[s] + r, i = perm[p];
  long long off = so.off[s];
  double *Wb = so.W + 9*off + l;
  int *cb = so.wcol + off + l;
  for (int k = g, m = 0; k < si.rowptr[i + 1] - si.rowptr[i]; k += G, m ++)
  {
    int q = si.rowptr[i] + k;
#pragma unroll
    for (int e = 0; e < 9; e ++) Wb[(9*(size_t)m + e)*SLICE] = si.W[9*(size_t)q + e];
    cb[(size_t)m*SLICE] = iperm[si.col[q]];
  }
}

// Workgroup b of a launch runs on XCD b mod 8.  With xcd set, the blocks of
// one XCD take one contiguous run of the colour's slices, so the rows an XCD
// updates are neighbours in the row order, and (with a spatial row order)
// their neighbours' reactions are shared in that XCD's L2.
__device__ __forceinline__ int xcd_block(int b, int nb)
{
  const int x = b & 7, k = b >> 3, q = nb >> 3, rm = nb & 7;
  return x*q + min(x, rm) + k;
}

template <int NT = TPB>
__device__ void finish_sweep(Ctrl *ctrl, const double *partials, int np, double *hist);

// fin (the last launch of a sweep only): the block that completes the launch
// last takes the sweep's termination test, as k_finish would after it
// WPB wavefronts per workgroup (one slice each)
template <typename T, int G, bool MERIT, int WPB = TPB/SLICE>
__global__ __launch_bounds__(WPB*64) void k_sweep(SweepArgs<T> a, int s0, int s1, int pbase, int upd_lo, int upd_hi,
                                                   int backward, int xcd, int *fin, int np, double *hist)
{
  // uniform: set only between launches (by the finish of an earlier sweep); read
  // with the slice table, not before it
  const int done = a.ctrl->done;
  const int bb = xcd ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
  const int s = s0 + bb*WPB + (threadIdx.x >> 6);
  double eu = 0.0, el = 0.0, up = 0.0;
  sweep_slice<T, G, MERIT>(a, s, s < s1, upd_lo, upd_hi, backward, &a.ctrl->sweeps, a.failkey, a.failcount,
                           &a.ctrl->aborted, eu, el, up, done == 0);
  if (done) return;
  block_sum3<WPB*64>(eu, el, up);
  if (threadIdx.x == 0)
  {
    double *pt = a.partials + 3*(size_t)(pbase + blockIdx.x);
    pt[0] = eu; pt[1] = el; pt[2] = up;
  }
  if (fin)
  {
    __shared__ int last;
    if (threadIdx.x == 0)
    {
      __threadfence();   // this block's partial before its ticket
      last = atomicAdd(fin, 1) == (int)gridDim.x - 1;
    }
    __syncthreads();
    if (last)
    {
      __threadfence();   // every other block's partial is visible
      finish_sweep<WPB*64>(a.ctrl, a.partials, np, hist);
      if (threadIdx.x == 0) *fin = 0;
    }
  }
}

// Several consecutive colours of a sweep, each of at most a few slices, in one
// launch of one workgroup: colour after colour in sweep order, wavefront w
// taking slices w, w + NW, ... of each, with a workgroup barrier between
// colours.  The rows of a colour then read the reactions the colour before
// wrote through the workgroup's own L1, which the barrier's workgroup-scope
// fence makes visible: the same iterates as one launch per colour, one launch
// fewer per colour merged.  One partial (3 sums) for the group; fin as in
// k_sweep (the group's block is the launch's last).
struct ColourGroup {
  static const int MAX = 8;
  int n;
  int s0[MAX], s1[MAX], lo[MAX], hi[MAX];
};

template <typename T, int G, bool MERIT, int NW>
__global__ __launch_bounds__(NW*64) void k_sweep_group(SweepArgs<T> a, ColourGroup cg, int pbase, int backward,
                                                        int *fin, int np, double *hist)
{
  const int done = a.ctrl->done;
  const int w = threadIdx.x >> 6;
  double eu = 0.0, el = 0.0, up = 0.0;
  for (int q = 0; q < cg.n; q ++)
  {
    for (int s = cg.s0[q] + w; s < cg.s1[q]; s += NW)
      sweep_slice<T, G, MERIT>(a, s, true, cg.lo[q], cg.hi[q], backward, &a.ctrl->sweeps, a.failkey, a.failcount,
                               &a.ctrl->aborted, eu, el, up, done == 0);
    __syncthreads();   // this colour's reactions before the next colour's gathers
  }
  if (done) return;
  block_sum3<NW*64>(eu, el, up);
  if (threadIdx.x == 0)
  {
    double *pt = a.partials + 3*(size_t)pbase;
    pt[0] = eu; pt[1] = el; pt[2] = up;
  }
  if (fin)
  {
    __syncthreads();
    if (threadIdx.x == 0) __threadfence();   // the partial before the sums
    __syncthreads();
    finish_sweep<NW*64>(a.ctrl, a.partials, np, hist);
  }
}

template <int G>
__global__ __launch_bounds__(TPB) void k_merit(MeritArgs a, int nslices)
{
  const int s = blockIdx.x*(TPB/SLICE) + (threadIdx.x >> 6);
  double up = 0.0, z1 = 0.0, z2 = 0.0;
  merit_slice<G>(a, s, s < nslices, up);
  block_sum3(up, z1, z2);
  if (threadIdx.x == 0) a.mpartials[blockIdx.x] = up;
}

// end of a sweep (bgs.c:1021-1036): errors, merit, histories, termination.
// Without merit, sweep k closes iteration k.  With merit, sweep k >= 1 closes
// iteration k-1 (its merit was formed during sweep k); sweep 0 only records its
// error.  ctrl->final_buf / final_U name the buffers holding the final state.
template <int NT>
__device__ void finish_sweep(Ctrl *ctrl, const double *partials, int np, double *hist)
{
  double eu = 0.0, el = 0.0, up = 0.0;
  for (int k = threadIdx.x; k < np; k += NT)
  {
    eu += partials[3*(size_t)k]; el += partials[3*(size_t)k + 1]; up += partials[3*(size_t)k + 2];
  }
  block_sum3<NT>(eu, el, up);
  if (threadIdx.x != 0) return;
  const int k = ctrl->sweeps;
  const double error_k = sqrt(eu) / sqrt(el == 0.0 ? 1.0 : el);
  ctrl->sweeps = k + 1;
  if (ctrl->nomerit)
  {
    hist[k] = error_k;
    hist[ctrl->maxiter + k] = 0.0;
    ctrl->iters = k + 1;
    ctrl->error = error_k;
    ctrl->merit = 0.0;
    ctrl->final_buf = (k + 1) & 1;
    ctrl->final_U = 0;
    if (ctrl->aborted || !(k + 1 < ctrl->maxiter && (error_k > ctrl->eps || k + 1 < ctrl->miniter))) ctrl->done = 1;
    return;
  }
  if (ctrl->aborted)
  {
    ctrl->iters = k + 1;
    ctrl->final_buf = (k + 1) & 1;
    ctrl->final_U = 0;
    ctrl->done = 1;
    return;
  }
  if (k == 0) { ctrl->prev_error = error_k; return; }
  const double lo = ctrl->energy < DBL_EPSILON ? 1.0 : ctrl->energy;
  const double merit = (0.5*up) / lo, error = ctrl->prev_error;
  hist[k - 1] = error;
  hist[ctrl->maxiter + k - 1] = merit;
  ctrl->iters = k;
  ctrl->error = error;
  ctrl->merit = merit;
  if (!(k < ctrl->maxiter && (error > ctrl->eps || merit > ctrl->meritval || k < ctrl->miniter)))
  {
    ctrl->final_buf = k & 1;   // P of sweep k = R after sweep k-1
    ctrl->final_U = 1;         // U after the merit of iteration k-1
    ctrl->done = 1;
  }
  else ctrl->prev_error = error_k;
}

__global__ __launch_bounds__(TPB) void k_finish(Ctrl *ctrl, const double *partials, int np, double *hist)
{
  if (ctrl->done) return;
  finish_sweep(ctrl, partials, np, hist);
}

// learned warm-start features of row p: gap, B (3), W_ii (9), R_prev (3), row
// length, friction; written at row p (dst_nat = 0) or at the natural index
__global__ void k_features(int n, const int *perm, const double *gap, const double *B, const double *Wd,
                           const double *R, const int *rowlen, const double *fri, int natural, float *out)
{
  int p = blockIdx.x*blockDim.x + threadIdx.x;
  if (p >= n) return;
  int i = perm[p], o = natural ? i : p;
  float *f = out + 18*(size_t)o;
  f[0] = (float)gap[i];
  for (int k = 0; k < 3; k ++) f[1 + k] = (float)B[(size_t)k*n + p];
  for (int e = 0; e < 9; e ++) f[4 + e] = (float)Wd[(size_t)e*n + p];
  for (int k = 0; k < 3; k ++) f[13 + k] = (float)R[4*(size_t)p + k];
  f[16] = (float)rowlen[p];
  f[17] = (float)fri[p];
}

__global__ void k_to_f32(long long n, const double *src, float *dst)
{
  long long i = (long long)blockIdx.x*blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (float)src[i];
}

// FP32 row record of row p (colour-major), five float4: (B, rho), (friction,
// cohesion, shift, 2 rowlen + separated as integer bits), then W_ii and
// A = W_ii^-1 by their upper triangles, W00 W01 W02 W11 | W12 W22 A00 A01 |
// A02 A11 A12 A22 (column-major indices 0 3 6 4 7 8).  Both matrices are
// symmetric; their two triangles differ by far less than the FP32 rounding.
// The FP32 sweep reads its row with five 16-byte loads instead of 30 scalar ones.
__global__ void k_rec32(int n, const double *Wd, const double *A, const double *B, const double *rho,
                        const double *fri, const double *coh, const double *shift, const unsigned char *sep,
                        const int *rowlen, float4 *rec)
{
  int p = blockIdx.x*blockDim.x + threadIdx.x;
  if (p >= n) return;
  const size_t m = (size_t)n;
  float4 *r = rec + 5*(size_t)p;
  r[0] = make_float4((float)B[p], (float)B[m + p], (float)B[2*m + p], (float)rho[p]);
  r[1] = make_float4((float)fri[p], (float)coh[p], (float)shift[p], __int_as_float(2*rowlen[p] + (sep[p] ? 1 : 0)));
  r[2] = make_float4((float)Wd[p], (float)Wd[3*m + p], (float)Wd[6*m + p], (float)Wd[4*m + p]);
  r[3] = make_float4((float)Wd[7*m + p], (float)Wd[8*m + p], (float)A[p], (float)A[3*m + p]);
  r[4] = make_float4((float)A[6*m + p], (float)A[4*m + p], (float)A[7*m + p], (float)A[8*m + p]);
}

__global__ void k_to_f64(long long n, const float *src, double *dst)
{
  long long i = (long long)blockIdx.x*blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (double)src[i];
}

// =====================================================================
// results back to natural order, body update (bod.c:418-435, 1446-1447)
// =====================================================================

__global__ void k_unperm(int n, const int *perm, const double *R, const double *U, double *Rn, double *Un)
{
  int p = blockIdx.x*blockDim.x + threadIdx.x;
  if (p >= n) return;
  int i = perm[p];
#pragma unroll
  for (int k = 0; k < 3; k ++)
  {
    Rn[3*(size_t)i + k] = R[4*(size_t)p + k];
    Un[3*(size_t)i + k] = U[(size_t)k*n + p];
  }
}

__global__ void k_body_update(BodyView bv, ConView cv, int nb, const int *bptr, const int *blist,
                              const int *iperm, const double *R, double step, double *velo_out, double *force_out)
{
  int b = blockIdx.x*blockDim.x + threadIdx.x;
  if (b >= nb) return;
  double u[6];
#pragma unroll
  for (int k = 0; k < 6; k ++) u[k] = bv.velo[6*(size_t)b + k];
  if (bv.kind[b] == 0)
  {
    double f[6] = {0, 0, 0, 0, 0, 0};
    const double *Rot = bv.rot + 9*(size_t)b, *X0 = bv.x0 + 3*(size_t)b;
    for (int q = bptr[b]; q < bptr[b + 1]; q ++)
    {
      int i = blist[q];
      double sg = cv.master[i] == b ? 1.0 : -1.0, bs[9], Hr[9];
      const double *X = sg > 0.0 ? cv.mpnt + 3*(size_t)i : cv.spnt + 3*(size_t)i;
#pragma unroll
      for (int e = 0; e < 9; e ++) bs[e] = cv.base[9*(size_t)i + e];
      operator_Hr(Rot, X0, X, bs, Hr);
      int p = iperm[i];
      double r0 = R[4*(size_t)p], r1 = R[4*(size_t)p + 1], r2 = R[4*(size_t)p + 2];
#pragma unroll
      for (int c = 0; c < 3; c ++)
      {
        f[c] += sg * (Hr[3*c]*r0 + Hr[3*c + 1]*r1 + Hr[3*c + 2]*r2);
        f[3 + c] += sg * (bs[c]*r0 + bs[c + 3]*r1 + bs[c + 6]*r2);
      }
    }
    const double *J = bv.jinv + 9*(size_t)b;
    double minv = bv.minv[b];
#pragma unroll
    for (int r = 0; r < 3; r ++)
    {
      u[r] += step * (J[r]*f[0] + J[r + 3]*f[1] + J[r + 6]*f[2]);
      u[3 + r] += step * minv * f[3 + r];
    }
#pragma unroll
    for (int k = 0; k < 6; k ++) force_out[6*(size_t)b + k] = f[k];
  }
  else
  {
#pragma unroll
    for (int k = 0; k < 6; k ++) force_out[6*(size_t)b + k] = 0.0;
  }
#pragma unroll
  for (int k = 0; k < 6; k ++) velo_out[6*(size_t)b + k] = u[k];
}

// body-space products of NEWTON_SOLVER with locdyn OFF (U_WR_B, nts.c:396-431):
// H_trans_R and BODY_Invvec per body, v_b = h M_b^-1 sum_c (+/-) H_c' x_c with
// x_c at x + stride p, p = iperm[c]; obstacles carry none (bod.c:431)
__global__ void k_bs_bodies(BodyView bv, ConView cv, int nb, const int *bptr, const int *blist, const int *iperm,
                            const double *x, int stride, double step, double *v)
{
  int b = blockIdx.x*blockDim.x + threadIdx.x;
  if (b >= nb) return;
  double f[6] = {0, 0, 0, 0, 0, 0};
  if (bv.kind[b] == 0)
  {
    const double *Rot = bv.rot + 9*(size_t)b, *X0 = bv.x0 + 3*(size_t)b;
    for (int q = bptr[b]; q < bptr[b + 1]; q ++)
    {
      int i = blist[q];
      double sg = cv.master[i] == b ? 1.0 : -1.0, bs[9], Hr[9];
      const double *X = sg > 0.0 ? cv.mpnt + 3*(size_t)i : cv.spnt + 3*(size_t)i;
#pragma unroll
      for (int e = 0; e < 9; e ++) bs[e] = cv.base[9*(size_t)i + e];
      operator_Hr(Rot, X0, X, bs, Hr);
      const double *r = x + (size_t)stride*iperm[i];
      double r0 = r[0], r1 = r[1], r2 = r[2];
#pragma unroll
      for (int c = 0; c < 3; c ++)
      {
        f[c] += sg * (Hr[3*c]*r0 + Hr[3*c + 1]*r1 + Hr[3*c + 2]*r2);
        f[3 + c] += sg * (bs[c]*r0 + bs[c + 3]*r1 + bs[c + 6]*r2);
      }
    }
    const double *J = bv.jinv + 9*(size_t)b;
    const double minv = bv.minv[b];
    double u[6];
#pragma unroll
    for (int r = 0; r < 3; r ++)
    {
      u[r] = step * (J[r]*f[0] + J[r + 3]*f[1] + J[r + 6]*f[2]);
      u[3 + r] = step * minv * f[3 + r];
    }
#pragma unroll
    for (int k = 0; k < 6; k ++) f[k] = u[k];
  }
#pragma unroll
  for (int k = 0; k < 6; k ++) v[6*(size_t)b + k] = f[k];
}

// H_times_u per row p: z = H_m v_m - H_s v_s; mode 0 the step operator of
// Matvec (nts.c:584-652), y = beta y + alpha (X z + Y x_p) + alpha delta x_p;
// mode 1 U = B + z
__global__ __launch_bounds__(TPB) void k_bs_rows(BodyView bv, ConView cv, int n, const int *perm, const double *v,
                                                  int mode, const double *Xm, const double *Ym, const double *x,
                                                  double alpha, double beta, double delta, double *y,
                                                  const double *B, double *U)
{
  int p = blockIdx.x*blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int i = perm[p], m = cv.master[i], s = cv.slave[i];
  double bs[9], Hr[9], z[3], t[3];
#pragma unroll
  for (int k = 0; k < 9; k ++) bs[k] = cv.base[9*(size_t)i + k];
  operator_Hr(bv.rot + 9*(size_t)m, bv.x0 + 3*(size_t)m, cv.mpnt + 3*(size_t)i, bs, Hr);
  h_u(Hr, bs, v + 6*(size_t)m, z);
  if (s >= 0)
  {
    operator_Hr(bv.rot + 9*(size_t)s, bv.x0 + 3*(size_t)s, cv.spnt + 3*(size_t)i, bs, Hr);
    h_u(Hr, bs, v + 6*(size_t)s, t);
    z[0] -= t[0]; z[1] -= t[1]; z[2] -= t[2];
  }
  if (mode == 1)
  {
#pragma unroll
    for (int k = 0; k < 3; k ++) U[(size_t)k*n + p] = B[(size_t)k*n + p] + z[k];
    return;
  }
  if (mode == 2)
  {
#pragma unroll
    for (int k = 0; k < 3; k ++) y[3*(size_t)i + k] = z[k];   // natural order
    return;
  }
  const double xp[3] = {x[3*(size_t)p], x[3*(size_t)p + 1], x[3*(size_t)p + 2]};
  double X[9], Y[9];
#pragma unroll
  for (int e = 0; e < 9; e ++) { X[e] = Xm[(size_t)e*n + p]; Y[e] = Ym[(size_t)e*n + p]; }
  const double ad = alpha*delta;
#pragma unroll
  for (int k = 0; k < 3; k ++)
  {
    const double u = (X[k]*z[0] + X[k + 3]*z[1] + X[k + 6]*z[2]) + (Y[k]*xp[0] + Y[k + 3]*xp[1] + Y[k + 6]*xp[2]);
    double q = beta*y[3*(size_t)p + k];
    q = q + alpha*u;
    y[3*(size_t)p + k] = q + ad*xp[k];
  }
}

// 3 values per row from natural to permuted order
__global__ void k_perm3(int n, const int *perm, const double *xn, double *xp)
{
  const int p = blockIdx.x*blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int i = perm[p];
#pragma unroll
  for (int k = 0; k < 3; k ++) xp[3*(size_t)p + k] = xn[3*(size_t)i + k];
}

int sfgs::bodyspace_product(sfgs_ctx *ctx, const double *x, int stride, int mode, const double *X, const double *Y,
                            double alpha, double beta, double delta, double *y)
{
  const int n = ctx->nc, nb = ctx->nb;
  if (n == 0) return 0;
  if (!ctx->from_bodies) { ctx->err = "body-space products need a system assembled from bodies"; return SFGS_E_STATE; }
  CHK(ensure(ctx, ctx->bsv, 6*(size_t)std::max(nb, 1)));
  hipStream_t st = ctx->stream;
  BodyView bv{ctx->kind.p, ctx->rot.p, ctx->x0.p, ctx->jinv.p, ctx->minv.p, ctx->velo.p, ctx->velo0.p};
  ConView cv{ctx->master.p, ctx->slave.p, ctx->mpnt.p, ctx->spnt.p, ctx->base.p, ctx->gap.p, ctx->fri.p,
             ctx->rest.p, ctx->coh.p, ctx->Rn.p, ctx->Un.p};
  if (nb > 0)
    hipLaunchKernelGGL(k_bs_bodies, dim3(nblk(nb)), dim3(TPB), 0, st, bv, cv, nb, ctx->bptr.p, ctx->blist.p,
                       ctx->iperm.p, x, stride, ctx->step, ctx->bsv.p);
  hipLaunchKernelGGL(k_bs_rows, dim3(nblk(n)), dim3(TPB), 0, st, bv, cv, n, ctx->perm.p, ctx->bsv.p, mode, X, Y, x,
                     alpha, beta, delta, y, ctx->B.p, ctx->U.p);
  HIPCHK(hipGetLastError());
  return 0;
}

// =====================================================================
// penalty solver: the spring-dashpot law per contact (pes.c:33-106, 247-281)
// =====================================================================

// PENALTY_Spring_Dashpot_Contact (pes.c:33-106); the operations and their
// order are those of the reference, with no contraction into fused
// multiply-adds.  A singular tangential block leaves R_T as carried and U as it
// was (pes.c:72, the return value that PENALTY_Solve ignores).  coh is the
// contact's cohesion state (cohesive when > 0; it equals the material's
// cohesion times the area while the contact is cohesive): a cohesive contact
// takes tension down to -coh and loses its cohesion below it (mode I,
// pes.c:76-82) or on the cone (mode II, pes.c:88-99).
__device__ void penalty_law(int implicit, double h, double gap, double k, double c, double p, double mu,
                            const double *W, const double *B, const double *V, double *U, double *R, double *coh)
{
#pragma clang fp contract(off)
  const double BN = B[2] + W[2]*R[0] + W[5]*R[1];
  double g, s;
  if (implicit)
  {
    g = gap + 0.25*h*(BN - V[2]);
    g = g < 0.0 ? g : 0.0;
    s = k*p*pow(-g, p - 1.0);
    if (c < 0.0) c = 2.0*sqrt(h*s/W[8]);
    R[2] = (k*pow(-g, p) - 0.5*c*(BN + V[2])) / (1.0 + (0.25*h*s + 0.5*c)*W[8]);
  }
  else
  {
    g = gap < 0.0 ? gap : 0.0;
    s = k*p*pow(-g, p - 1.0);
    if (c < 0.0) c = 2.0*sqrt(h*s/W[8]);
    R[2] = k*pow(-g, p) - c*V[2];
  }
  bool cohesive = *coh > 0.0;
  if (!cohesive && R[2] < 0.0)
  {
    R[0] = R[1] = R[2] = 0.0;
    U[0] = U[1] = U[2] = 0.0;
    return;
  }
  const double BT0 = B[0] + W[6]*R[2], BT1 = B[1] + W[7]*R[2];
  const double det = W[0]*W[4] - W[1]*W[3];
  if (det == 0.0) return;
  const double I0 = W[4]/det, I1 = -W[1]/det, I2 = -W[3]/det, I3 = W[0]/det;
  R[0] = -I0*BT0 - I2*BT1;
  R[1] = -I1*BT0 - I3*BT1;
  if (cohesive && R[2] < -*coh)
  {
    cohesive = false;
    R[2] = -*coh;
    *coh = 0.0;
  }
  const double len = sqrt(R[0]*R[0] + R[1]*R[1]), lim = mu*fabs(R[2]);
  if (len > lim)
  {
    R[0] *= lim/len;
    R[1] *= lim/len;
    if (cohesive) *coh = 0.0;
  }
  for (int r = 0; r < 3; r ++) U[r] = B[r] + (W[r]*R[0] + W[r + 3]*R[1] + W[r + 6]*R[2]);
}

// spring, dashpot and Hertz power of each contact: given per contact, or from
// the material table (6 values per material: friction, restitution, cohesion,
// spring, dashpot, hpow) by the surfaces of the contact's bodies
struct PenMat {
  const double *given;    // 3 per contact, or null
  const int *surf, *mat_s;
  const double *mat_v, *def;
  int nmat;
};

__device__ void penalty_material(const PenMat &pm, const ConView &cv, int i, double *kcp)
{
  if (pm.given)
  {
    for (int e = 0; e < 3; e ++) kcp[e] = pm.given[3*(size_t)i + e];
    return;
  }
  const int sl = cv.slave[i];
  int s1 = pm.surf[cv.master[i]], s2 = sl >= 0 ? pm.surf[sl] : s1, a = s1 < s2 ? s1 : s2, b = s1 < s2 ? s2 : s1;
  const double *v = pm.def;
  for (int m = 0; m < pm.nmat; m ++)
    if (pm.mat_s[2*m] == a && pm.mat_s[2*m + 1] == b) { v = pm.mat_v + 6*m; break; }
  for (int e = 0; e < 3; e ++) kcp[e] = v[3 + e];
}

// PENALTY_Solve (pes.c:247-281) per contact i: W_ii, B, V as the assembly
// computes them (ldy.c:364-397, 654-716; the penalty update forms no
// off-diagonal blocks of contacts, ldy.c:744), then the law from the carried R.
// R goes to the 4-per-row store and U to the SoA planes in the identity order.
__global__ __launch_bounds__(TPB) void k_penalty(BodyView bv, ConView cv, PenMat pm, int n, int implicit,
                                                  double step, double *Rs, double *Us, double *coh)
{
  int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i >= n) return;
  int m = cv.master[i], s = cv.slave[i];
  double bs[9], Hr[9], Wi[9], P[9], y[3], Vv[3], Bv[3], R[3], U[3], kcp[3];
#pragma unroll
  for (int k = 0; k < 9; k ++) bs[k] = cv.base[9*(size_t)i + k];
  operator_Hr(bv.rot + 9*(size_t)m, bv.x0 + 3*(size_t)m, cv.mpnt + 3*(size_t)i, bs, Hr);
  h_u(Hr, bs, bv.velo0 + 6*(size_t)m, Vv);
  h_u(Hr, bs, bv.velo + 6*(size_t)m, Bv);
#pragma unroll
  for (int k = 0; k < 9; k ++) Wi[k] = 0.0;
  if (couples(bv.kind, m))
  {
    h_im_ht(Hr, bs, Hr, bs, bv.jinv + 9*(size_t)m, bv.minv[m], P);
#pragma unroll
    for (int k = 0; k < 9; k ++) Wi[k] += P[k];
  }
  if (s >= 0)
  {
    operator_Hr(bv.rot + 9*(size_t)s, bv.x0 + 3*(size_t)s, cv.spnt + 3*(size_t)i, bs, Hr);
    h_u(Hr, bs, bv.velo0 + 6*(size_t)s, y);
    Vv[0] -= y[0]; Vv[1] -= y[1]; Vv[2] -= y[2];
    h_u(Hr, bs, bv.velo + 6*(size_t)s, y);
    Bv[0] -= y[0]; Bv[1] -= y[1]; Bv[2] -= y[2];
    if (couples(bv.kind, s))
    {
      h_im_ht(Hr, bs, Hr, bs, bv.jinv + 9*(size_t)s, bv.minv[s], P);
#pragma unroll
      for (int k = 0; k < 9; k ++) Wi[k] += P[k];
    }
  }
#pragma unroll
  for (int k = 0; k < 9; k ++) Wi[k] *= step;
#pragma unroll
  for (int k = 0; k < 3; k ++) { R[k] = cv.Rn[3*(size_t)i + k]; U[k] = cv.Un[3*(size_t)i + k]; }
  penalty_material(pm, cv, i, kcp);
  double c = coh[i];
  penalty_law(implicit, step, cv.gap[i], kcp[0], kcp[1], kcp[2], cv.fri[i], Wi, Bv, Vv, U, R, &c);
  if (c != coh[i]) coh[i] = c;
#pragma unroll
  for (int k = 0; k < 3; k ++) { Rs[4*(size_t)i + k] = R[k]; Us[(size_t)k*n + i] = U[k]; }
  Rs[4*(size_t)i + 3] = 0.0;
}

// update_cohesion (ldy.c:69-95) per row p of the solved system, on the
// contact's cohesion state in natural order
__global__ void k_cohesion(int n, const int *perm, const double *R, const double *fri, double *coh)
{
#pragma clang fp contract(off)
  int p = blockIdx.x*blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int i = perm[p];
  const double c = coh[i];
  if (!(c > 0.0)) return;
  const double e = 1e-8*c, r0 = R[4*(size_t)p], r1 = R[4*(size_t)p + 1], r2 = R[4*(size_t)p + 2];
  if (r2 + c < e || sqrt(r0*r0 + r1*r1) + e >= fri[i]*(r2 + c)) coh[i] = 0.0;
}

__global__ void k_identity(int n, int *perm, int *iperm)
{
  int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i < n) { perm[i] = i; iperm[i] = i; }
}

// =====================================================================
// host side
// =====================================================================

static int scan_int(sfgs_ctx *ctx, const int *in, int *out, int n)
{
  size_t bytes = 0;
  HIPCHK(sfgs::scan_int(nullptr, bytes, in, out, n, ctx->stream));
  CHK(ensure(ctx, ctx->scratch, bytes));
  HIPCHK(sfgs::scan_int(ctx->scratch.p, bytes, in, out, n, ctx->stream));
  return 0;
}

static int scan_ll(sfgs_ctx *ctx, const long long *in, long long *out, int n)
{
  size_t bytes = 0;
  HIPCHK(sfgs::scan_ll(nullptr, bytes, in, out, n, ctx->stream));
  CHK(ensure(ctx, ctx->scratch, bytes));
  HIPCHK(sfgs::scan_ll(ctx->scratch.p, bytes, in, out, n, ctx->stream));
  return 0;
}

// colour the natural CSR (ctx->rowptr/col, n rows) into ctx->colour; sets ncolours
static int colour_graph(sfgs_ctx *ctx, int n)
{
  ctx->ncolours = 0;
  if (n == 0) return 0;
  CHK(ensure(ctx, ctx->colour, n));
  CHK(ensure(ctx, ctx->colour2, n));
  CHK(ensure(ctx, ctx->counter, 4));
  hipLaunchKernelGGL(k_fill_int, dim3(nblk(n)), dim3(TPB), 0, ctx->stream, n, -1, ctx->colour.p);
  hipLaunchKernelGGL(k_fill_int, dim3(nblk(n)), dim3(TPB), 0, ctx->stream, n, -1, ctx->colour2.p);
  HIPCHK(hipMemsetAsync(ctx->counter.p, 0, 4*sizeof(int), ctx->stream));
  int h[2] = {0, 0};
  for (int round = 0; ; round ++)
  {
    if (round > 100000) { ctx->err = "colouring did not finish"; return SFGS_E_STATE; }
    HIPCHK(hipMemsetAsync(ctx->counter.p, 0, sizeof(int), ctx->stream));
    hipLaunchKernelGGL(k_jp_round, dim3(nblk(n)), dim3(TPB), 0, ctx->stream, n, ctx->rowptr.p, ctx->col.p,
                       ctx->colour.p, ctx->colour2.p, ctx->counter.p, ctx->counter.p + 1);
    hipLaunchKernelGGL(k_copy_int, dim3(nblk(n)), dim3(TPB), 0, ctx->stream, n, ctx->colour2.p, ctx->colour.p);
    HIPCHK(hipMemcpyAsync(h, ctx->counter.p, 2*sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (h[0] == 0) break;
  }
  if (h[1] >= SFGS_MAX_COLOURS) { ctx->err = "contact graph needs more than 256 colours"; return SFGS_E_COLOURS; }
  ctx->ncolours = h[1] + 1;
  return 0;
}

// permutation into colour-major order and the SELL-64 slice table
static int build_order(sfgs_ctx *ctx, int n, bool have_master)
{
  const int G = ctx->G, RPS = SLICE / G;
  const bool scenes = ctx->nscenes > 1;
  ctx->G_built = G;
  CHK(ensure(ctx, ctx->skeys, n));
  CHK(ensure(ctx, ctx->skeys2, n));
  CHK(ensure(ctx, ctx->svals, n));
  CHK(ensure(ctx, ctx->perm, n));
  CHK(ensure(ctx, ctx->iperm, n));
  // orders 2 and 3 need the master bodies of the contacts (a system uploaded as W has none)
  const int mode = ctx->row_order >= 2 && !have_master ? 0 : ctx->row_order;
  int keybits = mode >= 2 ? 40 : 32;
  if (mode == 3)
  {
    CHK(ensure(ctx, ctx->bbox, 6));
    hipLaunchKernelGGL(k_bbox_init, dim3(1), dim3(64), 0, ctx->stream, ctx->bbox.p);
    if (ctx->nb > 0)
      hipLaunchKernelGGL(k_bbox, dim3(nblk(ctx->nb)), dim3(TPB), 0, ctx->stream, ctx->nb, ctx->kind.p, ctx->x0.p,
                         ctx->bbox.p);
  }
  if (scenes)
  {
    for (int v = ctx->nscenes - 1; v > 0; v >>= 1) keybits ++;
    CHK(ensure(ctx, ctx->row_scene, n));
    CHK(ensure(ctx, ctx->counter, 4));
    HIPCHK(hipMemsetAsync(ctx->counter.p, 0, sizeof(int), ctx->stream));
    int bad = 0;
    if (n > 0)
      hipLaunchKernelGGL(k_row_scene, dim3(nblk(n)), dim3(TPB), 0, ctx->stream, n, ctx->master.p, ctx->slave.p,
                         ctx->kind.p, ctx->body_scene.p, ctx->row_scene.p, ctx->counter.p);
    HIPCHK(hipMemcpyAsync(&bad, ctx->counter.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (bad) { ctx->err = std::to_string(bad) + " contacts couple bodies of different scenes"; return SFGS_E_ARG; }
  }
  if (n > 0)
  {
    hipLaunchKernelGGL(k_sort_keys, dim3(nblk(n)), dim3(TPB), 0, ctx->stream, n, ctx->colour.p, ctx->rowptr.p,
                       scenes ? ctx->row_scene.p : (const int*)nullptr, mode, ctx->master.p, ctx->x0.p,
                       mode == 3 ? ctx->bbox.p : (const unsigned long long *)nullptr, ctx->skeys.p, ctx->svals.p);
    size_t bytes = 0;
    HIPCHK(sfgs::sort_pairs_u64(nullptr, bytes, ctx->skeys.p, ctx->skeys2.p, ctx->svals.p, ctx->perm.p, n, keybits, ctx->stream));
    CHK(ensure(ctx, ctx->scratch, bytes));
    HIPCHK(sfgs::sort_pairs_u64(ctx->scratch.p, bytes, ctx->skeys.p, ctx->skeys2.p, ctx->svals.p, ctx->perm.p, n, keybits, ctx->stream));
    hipLaunchKernelGGL(k_iperm, dim3(nblk(n)), dim3(TPB), 0, ctx->stream, n, ctx->perm.p, ctx->iperm.p);
  }
  // rows per colour, or per (scene, colour) group
  const int ncol = ctx->ncolours, nc = scenes ? ctx->nscenes*ncol : ncol;
  CHK(ensure(ctx, ctx->counter, (size_t)std::max(nc, 4)));
  HIPCHK(hipMemsetAsync(ctx->counter.p, 0, sizeof(int)*(size_t)std::max(nc, 4), ctx->stream));
  if (n > 0 && !scenes)
    hipLaunchKernelGGL(k_colour_count, dim3(nblk(n, 1024)), dim3(1024), 0, ctx->stream, n, ctx->colour.p, ctx->counter.p, nc);
  if (n > 0 && scenes)
    hipLaunchKernelGGL(k_group_count, dim3(nblk(n)), dim3(TPB), 0, ctx->stream, n, ctx->colour.p, ctx->row_scene.p, ncol,
                       ctx->counter.p);
  std::vector<int> cnt((size_t)std::max(nc, 1), 0);
  if (nc > 0) HIPCHK(hipMemcpyAsync(cnt.data(), ctx->counter.p, sizeof(int)*(size_t)nc, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->h_colptr.assign((size_t)nc + 1, 0);
  ctx->h_cslice.assign((size_t)nc + 1, 0);
  std::vector<int> row0, nrows;
  for (int c = 0; c < nc; c ++)
  {
    ctx->h_colptr[c + 1] = ctx->h_colptr[c] + cnt[c];
    for (int r = ctx->h_colptr[c]; r < ctx->h_colptr[c + 1]; r += RPS)
    {
      row0.push_back(r);
      nrows.push_back(std::min(RPS, ctx->h_colptr[c + 1] - r));
    }
    ctx->h_cslice[c + 1] = (int)row0.size();
  }
  int ns = (int)row0.size();
  ctx->nslices = ns;
  CHK(upload(ctx, ctx->slice_row0, row0.data(), (size_t)ns));
  CHK(upload(ctx, ctx->slice_nrows, nrows.data(), (size_t)ns));
  CHK(ensure(ctx, ctx->slice_len, (size_t)ns + 1));
  DBuf<long long> &tmp = ctx->slice_off;  // reuse: lengths first, then offsets
  CHK(ensure(ctx, tmp, (size_t)ns + 1));
  DBuf<long long> lens;
  CHK(ensure(ctx, lens, (size_t)ns + 1));
  HIPCHK(hipMemsetAsync(lens.p, 0, sizeof(long long)*((size_t)ns + 1), ctx->stream));
  if (ns > 0)
    hipLaunchKernelGGL(k_slice_len, dim3(nblk(ns)), dim3(TPB), 0, ctx->stream, ns, ctx->slice_row0.p, ctx->slice_nrows.p, ctx->perm.p,
                       ctx->rowptr.p, G, ctx->slice_len.p, lens.p);
  CHK(scan_ll(ctx, lens.p, ctx->slice_off.p, ns + 1));
  long long total = 0;
  HIPCHK(hipMemcpyAsync(&total, ctx->slice_off.p + ns, sizeof(long long), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  release(lens);
  ctx->sell_blocks = total;
  CHK(ensure(ctx, ctx->W, (size_t)(9*total)));
  CHK(ensure(ctx, ctx->wcol, (size_t)total));
  if (total > 0)
  {
    HIPCHK(hipMemsetAsync(ctx->W.p, 0, sizeof(double)*(size_t)(9*total), ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->wcol.p, 0, sizeof(int)*(size_t)total, ctx->stream));
  }
  return 0;
}

static int ensure_rows(sfgs_ctx *ctx, int n)
{
  CHK(ensure(ctx, ctx->rowlen, n));
  CHK(ensure(ctx, ctx->Wd, 9*(size_t)n));
  CHK(ensure(ctx, ctx->A, 9*(size_t)n));
  CHK(ensure(ctx, ctx->B, 3*(size_t)n));
  CHK(ensure(ctx, ctx->V, 3*(size_t)n));
  CHK(ensure(ctx, ctx->rho, n));
  CHK(ensure(ctx, ctx->pfri, n));
  CHK(ensure(ctx, ctx->pcoh, n));
  CHK(ensure(ctx, ctx->shift, n));
  CHK(ensure(ctx, ctx->sep, n));
  CHK(ensure(ctx, ctx->R, 4*(size_t)n));
  CHK(ensure(ctx, ctx->Rb, 4*(size_t)n));
  CHK(ensure(ctx, ctx->U, 3*(size_t)n));
  CHK(ensure(ctx, ctx->Um, 3*(size_t)n));
  return 0;
}

// body -> contact incidence (bptr, blist), ascending contact index per body (stable sort)
static int build_incidence(sfgs_ctx *ctx)
{
  const int nc = ctx->nc, nb = ctx->nb;
  hipStream_t st = ctx->stream;
  CHK(ensure(ctx, ctx->bkeys, 2*(size_t)nc));
  CHK(ensure(ctx, ctx->bkeys2, 2*(size_t)nc));
  CHK(ensure(ctx, ctx->bvals, 2*(size_t)nc));
  CHK(ensure(ctx, ctx->blist, 2*(size_t)nc));
  CHK(ensure(ctx, ctx->bptr, (size_t)nb + 1));
  CHK(ensure(ctx, ctx->counter, (size_t)nb + 1));
  if (nc > 0)
  {
    hipLaunchKernelGGL(k_inc_keys, dim3(nblk(nc)), dim3(TPB), 0, st, nc, ctx->master.p, ctx->slave.p, ctx->kind.p,
                       ctx->bkeys.p, ctx->bvals.p);
    size_t bytes = 0;
    HIPCHK(sfgs::sort_pairs_i32(nullptr, bytes, ctx->bkeys.p, ctx->bkeys2.p, ctx->bvals.p, ctx->blist.p, 2*nc, 32, st));
    CHK(ensure(ctx, ctx->scratch, bytes));
    HIPCHK(sfgs::sort_pairs_i32(ctx->scratch.p, bytes, ctx->bkeys.p, ctx->bkeys2.p, ctx->bvals.p, ctx->blist.p, 2*nc, 32, st));
  }
  HIPCHK(hipMemsetAsync(ctx->counter.p, 0, sizeof(int)*((size_t)nb + 1), st));
  if (nc > 0) hipLaunchKernelGGL(k_inc_count, dim3(nblk(2*nc)), dim3(TPB), 0, st, 2*nc, ctx->bkeys2.p, ctx->counter.p);
  CHK(scan_int(ctx, ctx->counter.p, ctx->bptr.p, nb + 1));
  return 0;
}

#define COMMA ,
#define SFGS_DISPATCH_G(G, CALL) \
  switch (G) { \
  case 1: { constexpr int GG = 1; CALL; } break; \
  case 2: { constexpr int GG = 2; CALL; } break; \
  case 8: { constexpr int GG = 8; CALL; } break; \
  default: { constexpr int GG = 4; CALL; } break; }

static RowsOut rows_out(sfgs_ctx *ctx, int n, double *epart, int *singular)
{
  RowsOut o;
  o.n = n; o.rowlen = ctx->rowlen.p; o.Wd = ctx->Wd.p; o.A = ctx->A.p; o.B = ctx->B.p; o.V = ctx->V.p;
  o.rho = ctx->rho.p; o.pfri = ctx->pfri.p; o.pcoh = ctx->pcoh.p; o.shift = ctx->shift.p; o.R = ctx->R.p;
  o.U = ctx->U.p; o.sep = ctx->sep.p; o.energy_part = epart; o.singular = singular;
  return o;
}

static SellOut sell_out(sfgs_ctx *ctx)
{
  SellOut so;
  so.row0 = ctx->slice_row0.p; so.nrows = ctx->slice_nrows.p; so.off = ctx->slice_off.p;
  so.W = ctx->W.p; so.wcol = ctx->wcol.p;
  return so;
}

extern "C" {

const char* SFGS_Error_String(int code)
{
  switch (code)
  {
  case SFGS_OK: return "ok";
  case SFGS_E_HIP: return "HIP runtime error";
  case SFGS_E_ARG: return "invalid argument";
  case SFGS_E_STATE: return "call out of order";
  case SFGS_E_COLOURS: return "too many colours";
  case SFGS_E_SINGULAR: return "singular diagonal block";
  }
  return "unknown error";
}

} // extern "C"

// ---- colouring without allocation (SFGS_Colour_Static) ----
// The scratch of the colouring rounds lives in one file-scope device array:
// the next round's colours (SFGS_STATIC_COLOUR_MAX ints) and, at a 256-byte
// aligned offset after them, 4 counters.  Every device that runs the library
// holds its own copy until the process ends (4 MiB + 256 B).  In host code the
// array's name is not a device address: kernels get the address that
// hipGetSymbolAddress returns for the current device, kept per device.
static __device__ __align__(256) int scratch[SFGS_STATIC_COLOUR_MAX + 64];

static int *scratch_ptr[64];
static std::mutex scratch_mu;

static int* static_scratch_address(int device)
{
  if (device < 0 || device >= 64) return nullptr;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_ptr[device])
  {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) { (void)hipGetLastError(); return nullptr; }
    scratch_ptr[device] = static_cast<int*>(p);
  }
  return scratch_ptr[device];
}

// CSR check on the device: rowptr[0] = 0, non-decreasing, rowptr[n] = nnz
// (the length of col), columns in [0, n); col is read only inside [0, nnz).
// bad[0] counts the rows that fail (the colouring rounds read cin[col[p]])
__global__ void k_check_csr(int n, long long nnz, const int *rowptr, const int *col, int *bad)
{
  int v = blockIdx.x*blockDim.x + threadIdx.x;
  if (v >= n) return;
  long long a = rowptr[v], b = rowptr[v + 1];
  if (a < 0 || b < a || b > nnz || (v == 0 && a != 0) || (v == n - 1 && b != nnz)) { atomicAdd(bad, 1); return; }
  for (long long p = a; p < b; p ++)
    if (col[p] < 0 || col[p] >= n) { atomicAdd(bad, 1); return; }
}

extern "C" {

sfgs_ctx* SFGS_Create(int device)
{
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0 || device < 0 || device >= count) return nullptr;
  if (hipSetDevice(device) != hipSuccess) return nullptr;
  sfgs_ctx *ctx = new sfgs_ctx();
  ctx->device = device;
  // the scratch array of SFGS_Colour_Static on this device (a failure is
  // reported by that entry point, which then refuses to run)
  ctx->static_scratch = static_scratch_address(device);
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) { delete ctx; return nullptr; }
  (void)hipEventCreate(&ctx->ev0);
  (void)hipEventCreate(&ctx->ev1);
  (void)hipEventCreate(&ctx->ev2);
  (void)hipEventCreate(&ctx->ev3);
  return ctx;
}

void SFGS_Destroy(sfgs_ctx *ctx)
{
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  release(ctx->kind); release(ctx->rot); release(ctx->x0); release(ctx->jinv); release(ctx->minv);
  release(ctx->velo); release(ctx->velo0);
  release(ctx->master); release(ctx->slave); release(ctx->mpnt); release(ctx->spnt); release(ctx->base);
  release(ctx->gap); release(ctx->fri); release(ctx->rest); release(ctx->coh); release(ctx->Rn); release(ctx->Un);
  release(ctx->bptr); release(ctx->blist); release(ctx->bkeys); release(ctx->bkeys2); release(ctx->bvals); release(ctx->bvals2);
  release(ctx->rowptr); release(ctx->rowlen_n); release(ctx->col); release(ctx->via);
  release(ctx->colour); release(ctx->colour2); release(ctx->perm); release(ctx->iperm); release(ctx->counter);
  release(ctx->skeys); release(ctx->skeys2); release(ctx->svals); release(ctx->svals2); release(ctx->scratch); release(ctx->bbox);
  release(ctx->slice_row0); release(ctx->slice_nrows); release(ctx->slice_len); release(ctx->slice_off);
  release(ctx->rowlen); release(ctx->W); release(ctx->Wd); release(ctx->A); release(ctx->B); release(ctx->V);
  release(ctx->rho); release(ctx->pfri); release(ctx->pcoh); release(ctx->shift); release(ctx->R); release(ctx->U); release(ctx->pen); sfgs_newton_release(ctx); sfgs_bar_release(ctx);
  release(ctx->Rb); release(ctx->Um); release(ctx->sep); release(ctx->wcol);
  release(ctx->W32); release(ctx->Wd32); release(ctx->rec32); release(ctx->B32); release(ctx->R32); release(ctx->R32b); release(ctx->U32); release(ctx->Um32);
  release(ctx->rho32); release(ctx->pfri32); release(ctx->pcoh32); release(ctx->shift32);
  release(ctx->partials); release(ctx->mpartials); release(ctx->hist); release(ctx->failkey);
  release(ctx->failcount); release(ctx->ctrl); release(ctx->fin);
  release(ctx->body_scene); release(ctx->row_scene); release(ctx->bscene_tab); release(ctx->bwg_scene);
  release(ctx->bgptr); release(ctx->bgslice); release(ctx->bbar); release(ctx->bpart); release(ctx->bout);
  release(ctx->babort); release(ctx->flow_mem);
  release(ctx->u_clist); release(ctx->u_sscene); release(ctx->u_sbound); release(ctx->u_scene);
  release(ctx->u_active); release(ctx->u_spart);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->ev2) (void)hipEventDestroy(ctx->ev2);
  if (ctx->ev3) (void)hipEventDestroy(ctx->ev3);
  for (hipEvent_t e : ctx->prof_ev) (void)hipEventDestroy(e);
  release(ctx->body_out);
  release(ctx->bsv);
  release(ctx->feat);
  sfgs_mlp_release(ctx);
  sfgs_dom_release(ctx);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* SFGS_Last_Error(sfgs_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int SFGS_Bodies_Upload(sfgs_ctx *ctx, int nb, const int *kind, const double *rot, const double *x0,
                       const double *jinv, const double *minv, const double *velo, const double *velo0)
{
  if (!ctx || nb < 0 || (nb > 0 && (!kind || !rot || !x0 || !jinv || !minv || !velo || !velo0))) return SFGS_E_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  ctx->nb = nb;
  CHK(upload(ctx, ctx->kind, kind, (size_t)nb));
  CHK(upload(ctx, ctx->rot, rot, 9*(size_t)nb));
  CHK(upload(ctx, ctx->x0, x0, 3*(size_t)nb));
  CHK(upload(ctx, ctx->jinv, jinv, 9*(size_t)nb));
  CHK(upload(ctx, ctx->minv, minv, (size_t)nb));
  CHK(upload(ctx, ctx->velo, velo, 6*(size_t)nb));
  CHK(upload(ctx, ctx->velo0, velo0, 6*(size_t)nb));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->have_bodies = 1;
  ctx->assembled = 0;
  ctx->nscenes = 1;     // scene ids refer to one body upload
  ctx->dom = 0;         // so do the shapes of the device loop
  return 0;
}

int SFGS_Batch_Scenes(sfgs_ctx *ctx, int nscenes, const int *body_scene)
{
  if (!ctx || nscenes < 1 || (nscenes > 1 && !body_scene)) return SFGS_E_ARG;
  if (!ctx->have_bodies) { ctx->err = "bodies must be uploaded before their scenes"; return SFGS_E_STATE; }
  if (nscenes > 1)
  {
    for (int b = 0; b < ctx->nb; b ++)
      if (body_scene[b] < 0 || body_scene[b] >= nscenes)
      {
        ctx->err = "body " + std::to_string(b) + " has an invalid scene";
        return SFGS_E_ARG;
      }
    HIPCHK(hipSetDevice(ctx->device));
    CHK(upload(ctx, ctx->body_scene, body_scene, (size_t)ctx->nb));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  }
  ctx->nscenes = nscenes;
  ctx->assembled = 0;
  return 0;
}

int SFGS_Contacts_Upload(sfgs_ctx *ctx, int nc, const int *master, const int *slave,
                         const double *mpnt, const double *spnt, const double *base,
                         const double *gap, const double *friction, const double *restitution,
                         const double *cohesion, const double *R, const double *U)
{
  if (!ctx || nc < 0) return SFGS_E_ARG;
  if (nc > 0 && (!master || !slave || !mpnt || !spnt || !base || !gap || !friction || !restitution || !cohesion || !R || !U))
    return SFGS_E_ARG;
  if (!ctx->have_bodies) { ctx->err = "bodies must be uploaded before contacts"; return SFGS_E_STATE; }
  for (int i = 0; i < nc; i ++)
  {
    if (master[i] < 0 || master[i] >= ctx->nb || slave[i] < -1 || slave[i] >= ctx->nb || slave[i] == master[i])
    {
      ctx->err = "contact " + std::to_string(i) + " has an invalid body pair";
      return SFGS_E_ARG;
    }
  }
  HIPCHK(hipSetDevice(ctx->device));
  ctx->nc = nc;
  CHK(upload(ctx, ctx->master, master, (size_t)nc));
  CHK(upload(ctx, ctx->slave, slave, (size_t)nc));
  CHK(upload(ctx, ctx->mpnt, mpnt, 3*(size_t)nc));
  CHK(upload(ctx, ctx->spnt, spnt, 3*(size_t)nc));
  CHK(upload(ctx, ctx->base, base, 9*(size_t)nc));
  CHK(upload(ctx, ctx->gap, gap, (size_t)nc));
  CHK(upload(ctx, ctx->fri, friction, (size_t)nc));
  CHK(upload(ctx, ctx->rest, restitution, (size_t)nc));
  CHK(upload(ctx, ctx->coh, cohesion, (size_t)nc));
  CHK(upload(ctx, ctx->Rn, R, 3*(size_t)nc));
  CHK(upload(ctx, ctx->Un, U, 3*(size_t)nc));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->assembled = 0;
  return 0;
}

// LOCDYN_Update_Begin from the bodies: full = true is UPALL (diagonal and
// off-diagonal blocks, adjacency, colouring); full = false is UPMIN (ldy.c:59,
// 644): the diagonal blocks, B and V only, every row in one colour, no
// adjacency, for the solvers that form W products through the bodies
static int locdyn_update(sfgs_ctx *ctx, double step, int dynamic, bool full)
{
  if (!ctx || !(step > 0.0)) return SFGS_E_ARG;
  if (!ctx->have_bodies) { ctx->err = "no bodies"; return SFGS_E_STATE; }
  HIPCHK(hipSetDevice(ctx->device));
  const int nc = ctx->nc;
  hipStream_t st = ctx->stream;

  CHK(build_incidence(ctx));

  // contact adjacency CSR (natural order); empty rows for UPMIN
  CHK(ensure(ctx, ctx->rowlen_n, (size_t)nc + 1));
  CHK(ensure(ctx, ctx->rowptr, (size_t)nc + 1));
  HIPCHK(hipMemsetAsync(ctx->rowlen_n.p, 0, sizeof(int)*((size_t)nc + 1), st));
  if (nc > 0 && full)
    hipLaunchKernelGGL(k_rowlen, dim3(nblk(nc)), dim3(TPB), 0, st, nc, ctx->master.p, ctx->slave.p, ctx->kind.p,
                       ctx->bptr.p, ctx->rowlen_n.p);
  CHK(scan_int(ctx, ctx->rowlen_n.p, ctx->rowptr.p, nc + 1));
  int nnz = 0;
  HIPCHK(hipMemcpyAsync(&nnz, ctx->rowptr.p + nc, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  ctx->nnz = nnz;
  CHK(ensure(ctx, ctx->col, (size_t)nnz));
  CHK(ensure(ctx, ctx->via, (size_t)nnz));
  if (nc > 0 && nnz > 0)
    hipLaunchKernelGGL(k_adj_fill, dim3(nblk(nc)), dim3(TPB), 0, st, nc, ctx->master.p, ctx->slave.p, ctx->kind.p,
                       ctx->bptr.p, ctx->blist.p, ctx->rowptr.p, ctx->col.p, ctx->via.p);

  // colouring and colour-major layout
  CHK(colour_graph(ctx, nc));
  CHK(build_order(ctx, nc, true));
  CHK(ensure_rows(ctx, nc));

  // rows and blocks
  int nbk = nblk(nc);
  CHK(ensure(ctx, ctx->partials, 2*(size_t)std::max(nbk, 1)));
  CHK(ensure(ctx, ctx->failcount, 4));
  HIPCHK(hipMemsetAsync(ctx->failcount.p, 0, sizeof(int)*4, st));
  BodyView bv{ctx->kind.p, ctx->rot.p, ctx->x0.p, ctx->jinv.p, ctx->minv.p, ctx->velo.p, ctx->velo0.p};
  ConView cv{ctx->master.p, ctx->slave.p, ctx->mpnt.p, ctx->spnt.p, ctx->base.p, ctx->gap.p, ctx->fri.p,
             ctx->rest.p, ctx->coh.p, ctx->Rn.p, ctx->Un.p};
  if (nc > 0)
  {
    hipLaunchKernelGGL(k_rows, dim3(nbk), dim3(TPB), 0, st, bv, cv, rows_out(ctx, nc, ctx->partials.p, ctx->failcount.p),
                       ctx->perm.p, ctx->rowptr.p, step, dynamic);
    if (ctx->nslices > 0)
      SFGS_DISPATCH_G(ctx->G_built, auto kern = k_offdiag<GG>; hipLaunchKernelGGL(kern, dim3(nblk((long long)ctx->nslices*SLICE)), dim3(TPB),
                      0, st, bv, cv, sell_out(ctx), ctx->perm.p, ctx->iperm.p, ctx->rowptr.p, ctx->col.p, ctx->via.p,
                      ctx->nslices, step));
  }
  std::vector<double> ep((size_t)std::max(nbk, 1), 0.0);
  int singular = 0;
  if (nc > 0) HIPCHK(hipMemcpyAsync(ep.data(), ctx->partials.p, sizeof(double)*(size_t)nbk, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(&singular, ctx->failcount.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipGetLastError());
  double e = 0.0;
  for (int k = 0; k < nbk; k ++) e += ep[k];
  ctx->energy = 0.5 * e;
  ctx->step = step;
  ctx->dynamic = dynamic;
  ctx->assembled = 1;
  ctx->from_bodies = 1;
  ctx->diag_only = 0;
  ctx->upmin = full ? 0 : 1;
  ctx->mirror_valid = 0;
  if (singular) { ctx->err = std::to_string(singular) + " singular diagonal blocks"; return SFGS_E_SINGULAR; }
  return 0;
}

int SFGS_Locdyn_Update_Begin(sfgs_ctx *ctx, double step, int dynamic)
{
  return locdyn_update(ctx, step, dynamic, true);
}

int SFGS_Locdyn_Update_Min(sfgs_ctx *ctx, double step, int dynamic)
{
  return locdyn_update(ctx, step, dynamic, false);
}

int SFGS_System_Upload(sfgs_ctx *ctx, int n, const int *rowptr, const int *col, const double *W,
                       const double *Wd, const double *A, const double *B, const double *V,
                       const double *rho, const double *friction, const double *restitution,
                       const double *cohesion, const double *gap, double free_energy,
                       double step, int dynamic, const double *R, const double *U)
{
  if (!ctx || n < 0 || !(step > 0.0) || !rowptr) return SFGS_E_ARG;
  if (n > 0 && (!Wd || !A || !B || !V || !rho || !friction || !restitution || !cohesion || !gap || !R || !U)) return SFGS_E_ARG;
  long long nnz = rowptr[n];
  if (rowptr[0] != 0 || nnz < 0 || (nnz > 0 && (!col || !W))) return SFGS_E_ARG;
  for (int i = 0; i < n; i ++) if (rowptr[i + 1] < rowptr[i]) return SFGS_E_ARG;
  for (long long q = 0; q < nnz; q ++) if (col[q] < 0 || col[q] >= n || col[q] == -1) return SFGS_E_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  ctx->nc = n;
  ctx->nnz = nnz;
  CHK(upload(ctx, ctx->rowptr, rowptr, (size_t)n + 1));
  CHK(upload(ctx, ctx->col, col, (size_t)nnz));
  DBuf<double> dW, dWd, dA, dB, dV, drho, dfri, drest, dcoh, dgap;
  CHK(upload(ctx, dW, W, 9*(size_t)nnz));
  CHK(upload(ctx, dWd, Wd, 9*(size_t)n));
  CHK(upload(ctx, dA, A, 9*(size_t)n));
  CHK(upload(ctx, dB, B, 3*(size_t)n));
  CHK(upload(ctx, dV, V, 3*(size_t)n));
  CHK(upload(ctx, drho, rho, (size_t)n));
  CHK(upload(ctx, dfri, friction, (size_t)n));
  CHK(upload(ctx, drest, restitution, (size_t)n));
  CHK(upload(ctx, dcoh, cohesion, (size_t)n));
  CHK(upload(ctx, dgap, gap, (size_t)n));
  CHK(upload(ctx, ctx->Rn, R, 3*(size_t)n));
  CHK(upload(ctx, ctx->Un, U, 3*(size_t)n));
  int rc = colour_graph(ctx, n);
  if (!rc) rc = build_order(ctx, n, false);
  if (!rc) rc = ensure_rows(ctx, n);
  if (!rc && n > 0)
  {
    SysIn si{ctx->rowptr.p, ctx->col.p, dW.p, dWd.p, dA.p, dB.p, dV.p, drho.p, dfri.p, drest.p, dcoh.p, dgap.p,
             ctx->Rn.p, ctx->Un.p};
    hipLaunchKernelGGL(k_rows_sys, dim3(nblk(n)), dim3(TPB), 0, st, si, rows_out(ctx, n, nullptr, nullptr),
                       ctx->perm.p, step, dynamic);
    if (ctx->nslices > 0)
      SFGS_DISPATCH_G(ctx->G_built, auto kern = k_offdiag_sys<GG>; hipLaunchKernelGGL(kern, dim3(nblk((long long)ctx->nslices*SLICE)),
                      dim3(TPB), 0, st, si, sell_out(ctx), ctx->perm.p, ctx->iperm.p, ctx->nslices));
  }
  hipError_t e = hipStreamSynchronize(st);
  release(dW); release(dWd); release(dA); release(dB); release(dV); release(drho); release(dfri);
  release(drest); release(dcoh); release(dgap);
  if (rc) return rc;
  if (e != hipSuccess) { ctx->err = hipGetErrorString(e); return SFGS_E_HIP; }
  ctx->energy = free_energy;
  ctx->step = step;
  ctx->dynamic = dynamic;
  ctx->assembled = 1;
  ctx->from_bodies = 0;
  ctx->diag_only = 0;
  ctx->upmin = 0;
  ctx->mirror_valid = 0;
  return 0;
}

} // extern "C"

// one sweep over all colours (forward or backward) with or without the fused merit;
// sweep k reads buffer X[k & 1] and writes X[(k + 1) & 1].  The last launch also
// takes the termination test (fin); returns false when nothing was launched
// (then k_finish takes it)
// one colour launch of k_sweep with the context's lanes per row and wavefronts per workgroup
#define SFGS_SWEEP_WPB(MER, GG) \
  switch (ctx->wpb_run) { \
  case 1: hipLaunchKernelGGL((k_sweep<T, GG, MER, 1>), dim3(grid[c]), dim3(64), 0, ctx->stream, a, ctx->h_cslice[c], \
                             ctx->h_cslice[c + 1], pbase[c], lo, hi, backward, xcd, f, np, hist); break; \
  case 2: hipLaunchKernelGGL((k_sweep<T, GG, MER, 2>), dim3(grid[c]), dim3(128), 0, ctx->stream, a, ctx->h_cslice[c], \
                             ctx->h_cslice[c + 1], pbase[c], lo, hi, backward, xcd, f, np, hist); break; \
  default: hipLaunchKernelGGL((k_sweep<T, GG, MER, 4>), dim3(grid[c]), dim3(256), 0, ctx->stream, a, ctx->h_cslice[c], \
                              ctx->h_cslice[c + 1], pbase[c], lo, hi, backward, xcd, f, np, hist); break; }
#define SFGS_DISPATCH_SWEEP(MER) \
  switch (ctx->G_built) { \
  case 1: SFGS_SWEEP_WPB(MER, 1) break; \
  case 2: SFGS_SWEEP_WPB(MER, 2) break; \
  case 8: SFGS_SWEEP_WPB(MER, 8) break; \
  default: SFGS_SWEEP_WPB(MER, 4) break; }

// the trailing colours merged into one k_sweep_group launch: wavefronts of its
// workgroup, and the most slices a colour may have to join it (one slice per
// wavefront)
static const int GROUP_NW = 8;

#define SFGS_GROUP_LAUNCH(MER, GG) \
  hipLaunchKernelGGL((k_sweep_group<T, GG, MER, GROUP_NW>), dim3(1), dim3(GROUP_NW*64), 0, ctx->stream, a, cg, \
                     gslot, backward, f, np, hist);
#define SFGS_DISPATCH_GROUP(MER) \
  switch (ctx->G_built) { \
  case 1: SFGS_GROUP_LAUNCH(MER, 1) break; \
  case 2: SFGS_GROUP_LAUNCH(MER, 2) break; \
  case 8: SFGS_GROUP_LAUNCH(MER, 8) break; \
  default: SFGS_GROUP_LAUNCH(MER, 4) break; }

// one sweep: a launch per colour c < gc in sweep order, and the colours
// gc .. ncol-1 (at most GROUP_NW slices each) in one k_sweep_group launch,
// last in a forward sweep and first in a backward one; gslot is the group's
// partial.  The last launch takes the termination test (fin)
template <typename T>
static bool launch_sweep(sfgs_ctx *ctx, const SweepArgs<T> &base, T *X0, T *X1, int k, int backward, bool merit,
                         const std::vector<int> &grid, const std::vector<int> &pbase, int gc, int gslot, bool prof,
                         int *fin, int np, double *hist)
{
  const int ncol = ctx->ncolours, n = ctx->nc;
  const int xcd = ctx->schedule == 2 ? 0 : 1;
  SweepArgs<T> a = base;
  if (k & 1) { a.P = X1; a.Q = X0; }
  else { a.P = X0; a.Q = X1; }
  // launch units in sweep order: a colour, or -1 for the group
  int units[SFGS_MAX_COLOURS + 1], nu = 0;
  bool grouped = false;
  for (int ci = 0; ci < ncol; ci ++)
  {
    int c = backward ? ncol - 1 - ci : ci;
    if (c >= gc) { if (!grouped) { units[nu ++] = -1; grouped = true; } continue; }
    if (grid[c] > 0) units[nu ++] = c;
  }
  for (int u = 0; u < nu; u ++)
  {
    int *f = u == nu - 1 ? fin : nullptr;
    const int c = units[u];
    if (c < 0)
    {
      ColourGroup cg;
      cg.n = 0;
      for (int ci = 0; ci < ncol; ci ++)
      {
        int cc = backward ? ncol - 1 - ci : ci;
        if (cc < gc || ctx->h_cslice[cc + 1] == ctx->h_cslice[cc]) continue;
        cg.s0[cg.n] = ctx->h_cslice[cc];
        cg.s1[cg.n] = ctx->h_cslice[cc + 1];
        cg.lo[cg.n] = backward ? ctx->h_colptr[cc + 1] : 0;
        cg.hi[cg.n] = backward ? n : ctx->h_colptr[cc];
        cg.n ++;
      }
      if (merit)
      {
        SFGS_DISPATCH_GROUP(true);
      }
      else
      {
        SFGS_DISPATCH_GROUP(false);
      }
      continue;
    }
    int lo = backward ? ctx->h_colptr[c + 1] : 0, hi = backward ? n : ctx->h_colptr[c];
    if (prof) (void)hipEventRecord(ctx->prof_ev[2*c], ctx->stream);
    if (merit)
    {
      SFGS_DISPATCH_SWEEP(true);
    }
    else
    {
      SFGS_DISPATCH_SWEEP(false);
    }
    if (prof) (void)hipEventRecord(ctx->prof_ev[2*c + 1], ctx->stream);
  }
  return nu > 0;
}

struct GsPlan {
  std::vector<int> grid, pbase;
  int np = 0;
  int gc = 0, gslot = 0;   // first colour of the merged trailing group (ncolours: none), its partial
  bool prof = false;
  double sweep_ms = 0.0;
  long long launches = 0;
};

// Gauss-Seidel iterations until k_finish stops them; hc receives the final control block
template <typename T>
static int run_gs(sfgs_ctx *ctx, SweepArgs<T> sa, T *X0, T *X1, int maxiter, int nomerit, double eps,
                  double meritval, int reverse, double *hist, GsPlan &pl, Ctrl &hc, int miniter = 0)
{
  hipStream_t st = ctx->stream;
  const int ncol = ctx->ncolours;
  std::memset(&hc, 0, sizeof(hc));
  hc.eps = eps; hc.meritval = meritval; hc.maxiter = maxiter; hc.nomerit = nomerit; hc.energy = ctx->energy;
  hc.miniter = miniter;
  HIPCHK(hipMemcpyAsync(ctx->ctrl.p, &hc, sizeof(Ctrl), hipMemcpyHostToDevice, st));
  const int maxsweeps = nomerit ? maxiter : maxiter + 1;   // with merit, sweep k closes iteration k-1
  // With merit the final U is the merit pass's unless FAILURE_EXIT stops a
  // sweep.  The merit pass's U is not stored either: the caller forms it from
  // the final R (final_U = 1: k_merit with update_U)
  sa.store_U = sa.failure == SFGS_FAILURE_EXIT;
  sa.store_Um = 0;
  int batch = pl.prof ? 1 : 2, k = 0;
  while (k < maxsweeps)
  {
    int nbt = std::min(batch, maxsweeps - k);
    for (int j = 0; j < nbt; j ++)
    {
      int kk = k + j;
      if (!launch_sweep<T>(ctx, sa, X0, X1, kk, reverse && (kk % 2), !nomerit && kk > 0, pl.grid, pl.pbase, pl.gc,
                           pl.gslot, pl.prof, ctx->fin.p, pl.np, hist))
        hipLaunchKernelGGL(k_finish, dim3(1), dim3(TPB), 0, st, ctx->ctrl.p, ctx->partials.p, pl.np, hist);
    }
    k += nbt;
    HIPCHK(hipMemcpyAsync(&hc, ctx->ctrl.p, sizeof(Ctrl), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    HIPCHK(hipGetLastError());
    if (pl.prof)
    {
      for (int ci = 0; ci < ncol; ci ++)
      {
        float e = 0.0f;
        HIPCHK(hipEventElapsedTime(&e, ctx->prof_ev[2*ci], ctx->prof_ev[2*ci + 1]));
        pl.sweep_ms += e;
        pl.launches ++;
      }
    }
    if (hc.done) break;
    batch = pl.prof ? 1 : std::min(batch*2, 32);
  }
  if (!hc.done) { hc.final_buf = maxsweeps & 1; hc.final_U = 0; }
  return 0;
}

// FP32 copies of the sweep data (built once per assembled system)
int sfgs::build_f32_mirror(sfgs_ctx *ctx)
{
  if (ctx->mirror_valid) return 0;
  const long long n = ctx->nc, nb = 9*ctx->sell_blocks;
  hipStream_t st = ctx->stream;
  CHK(ensure(ctx, ctx->W32, (size_t)std::max(nb, 1LL)));
  CHK(ensure(ctx, ctx->Wd32, 9*(size_t)n));
  CHK(ensure(ctx, ctx->rec32, 20*(size_t)n));
  CHK(ensure(ctx, ctx->B32, 3*(size_t)n));
  CHK(ensure(ctx, ctx->rho32, (size_t)n));
  CHK(ensure(ctx, ctx->pfri32, (size_t)n));
  CHK(ensure(ctx, ctx->pcoh32, (size_t)n));
  CHK(ensure(ctx, ctx->shift32, (size_t)n));
  CHK(ensure(ctx, ctx->R32, 4*(size_t)n));
  CHK(ensure(ctx, ctx->R32b, 4*(size_t)n));
  CHK(ensure(ctx, ctx->U32, 3*(size_t)n));
  CHK(ensure(ctx, ctx->Um32, 3*(size_t)n));
  if (nb > 0) hipLaunchKernelGGL(k_to_f32, dim3(nblk_ll(nb)), dim3(TPB), 0, st, nb, ctx->W.p, ctx->W32.p);
  if (n > 0)
  {
    hipLaunchKernelGGL(k_to_f32, dim3(nblk_ll(9*n)), dim3(TPB), 0, st, 9*n, ctx->Wd.p, ctx->Wd32.p);
    hipLaunchKernelGGL(k_rec32, dim3(nblk_ll(n)), dim3(TPB), 0, st, (int)n, ctx->Wd.p, ctx->A.p, ctx->B.p, ctx->rho.p,
                       ctx->pfri.p, ctx->pcoh.p, ctx->shift.p, ctx->sep.p, ctx->rowlen.p,
                       reinterpret_cast<float4 *>(ctx->rec32.p));
    hipLaunchKernelGGL(k_to_f32, dim3(nblk_ll(3*n)), dim3(TPB), 0, st, 3*n, ctx->B.p, ctx->B32.p);
    hipLaunchKernelGGL(k_to_f32, dim3(nblk_ll(n)), dim3(TPB), 0, st, n, ctx->rho.p, ctx->rho32.p);
    hipLaunchKernelGGL(k_to_f32, dim3(nblk_ll(n)), dim3(TPB), 0, st, n, ctx->pfri.p, ctx->pfri32.p);
    hipLaunchKernelGGL(k_to_f32, dim3(nblk_ll(n)), dim3(TPB), 0, st, n, ctx->pcoh.p, ctx->pcoh32.p);
    hipLaunchKernelGGL(k_to_f32, dim3(nblk_ll(n)), dim3(TPB), 0, st, n, ctx->shift.p, ctx->shift32.p);
  }
  ctx->mirror_valid = 1;
  return 0;
}

extern "C" {

int SFGS_Solve(sfgs_ctx *ctx, const SFGS_PARAM *par, double *rerhist, double *merhist, SFGS_REPORT *rep)
{
  if (!ctx || !par || par->maxiter <= 0 || !(par->epsilon > 0.0) || par->diagmaxiter <= 0) return SFGS_E_ARG;
  if (par->diagsolver < 0 || par->diagsolver > 3 || par->failure < 0 || par->failure > 2) return SFGS_E_ARG;
  if (par->precision != SFGS_PREC_FP64 && par->precision != SFGS_PREC_MIXED) return SFGS_E_ARG;
  if (par->precision == SFGS_PREC_MIXED && par->refine < 0) return SFGS_E_ARG;
  if (!ctx->assembled) { ctx->err = "solve before LOCDYN assembly"; return SFGS_E_STATE; }
  if (ctx->diag_only) { ctx->err = "the store holds a penalty solve, not an assembled system"; return SFGS_E_STATE; }
  if (ctx->upmin) { ctx->err = "the local dynamics hold the diagonal blocks only (SFGS_Locdyn_Update_Min)"; return SFGS_E_STATE; }
  if (ctx->nscenes > 1) { ctx->err = "the system holds independent scenes: use SFGS_Batch_Solve"; return SFGS_E_STATE; }
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int ncol = ctx->ncolours, maxiter = par->maxiter, n = ctx->nc;
  const int nomerit = par->meritval >= 1.0 ? 1 : 0;
  // MIXED: FP32 sweeps for at most maxiter - refine iterations, then FP64 sweeps
  // (at least refine of them) until the reference's termination test holds in
  // FP64 or the maxiter iterations are spent.  With maxiter <= refine the whole
  // solve is FP64.
  const bool mixed = par->precision == SFGS_PREC_MIXED && par->maxiter > par->refine;
  const int refine = mixed ? par->refine : 0;
  const int maxiter32 = maxiter - refine;
  // the dataflow schedule runs forward sweeps only
  const bool flow = ctx->schedule == 1 && !par->reverse && n > 0 && ctx->G_built <= 2;
  // the barrier schedule: one persistent launch, a device-wide barrier per colour
  const bool bar = ctx->schedule == 3 && n > 0 && ctx->G_built <= 2;

  // wavefronts per workgroup: the FP64 sweeps (239 VGPRs, two wavefronts per
  // SIMD) run faster in workgroups of 2, the FP32 ones in workgroups of 4
  ctx->wpb_run = ctx->wpb ? ctx->wpb : (mixed ? 4 : 2);
  GsPlan pl;
  pl.grid.assign((size_t)ncol, 0);
  pl.pbase.assign((size_t)ncol + 1, 0);
  for (int c = 0; c < ncol; c ++)
  {
    int ns = ctx->h_cslice[c + 1] - ctx->h_cslice[c];
    pl.grid[c] = (ns + ctx->wpb_run - 1) / ctx->wpb_run;
    pl.pbase[c + 1] = pl.pbase[c] + pl.grid[c];
  }
  pl.prof = par->profile != 0;
  // the trailing colours of at most GROUP_NW slices each (two or more of them,
  // at most ColourGroup::MAX) run as one launch; not in profile mode, which
  // times every colour
  pl.gc = ncol;
  if (!pl.prof)
  {
    int c = ncol, nonempty = 0;
    while (c > 0 && ncol - (c - 1) <= ColourGroup::MAX && ctx->h_cslice[c] - ctx->h_cslice[c - 1] <= GROUP_NW)
    {
      c --;
      if (ctx->h_cslice[c + 1] > ctx->h_cslice[c]) nonempty ++;
    }
    if (nonempty >= 2) pl.gc = c;
  }
  pl.gslot = pl.pbase[pl.gc];
  pl.np = pl.gc < ncol ? pl.gslot + 1 : pl.pbase[ncol];
  const int mgrid = (ctx->nslices + TPB/SLICE - 1) / (TPB/SLICE);
  CHK(ensure(ctx, ctx->partials, 3*(size_t)std::max(pl.np, 1)));
  CHK(ensure(ctx, ctx->mpartials, (size_t)std::max(mgrid, 1)));
  // histories: the first (or only) phase at hist (errors, then merits at
  // + its iteration bound), the FP64 phase of MIXED at hist + 2 maxiter
  CHK(ensure(ctx, ctx->hist, 4*(size_t)maxiter + 2));
  CHK(ensure(ctx, ctx->failkey, 1));
  CHK(ensure(ctx, ctx->failcount, 4));
  CHK(ensure(ctx, ctx->ctrl, 1));
  CHK(ensure(ctx, ctx->fin, 1));
  HIPCHK(hipMemsetAsync(ctx->fin.p, 0, sizeof(int), st));
  HIPCHK(hipMemsetAsync(ctx->partials.p, 0, sizeof(double)*3*(size_t)std::max(pl.np, 1), st));
  HIPCHK(hipMemsetAsync(ctx->failkey.p, 0, sizeof(unsigned long long), st));
  HIPCHK(hipMemsetAsync(ctx->failcount.p, 0, sizeof(int)*4, st));
  if (pl.prof && (int)ctx->prof_ev.size() < 2*ncol)
  {
    for (int k = (int)ctx->prof_ev.size(); k < 2*ncol; k ++)
    {
      hipEvent_t e;
      HIPCHK(hipEventCreate(&e));
      ctx->prof_ev.push_back(e);
    }
  }

  Ctrl hc, hr;
  std::memset(&hr, 0, sizeof(hr));
  // U = B + W R of the final R in ctx->R (MERIT_Function's update of U,
  // mrf.c:61-76), when the merit pass's U is the final one (final_U = 1)
  auto merit_U = [&]() {
    if (mgrid == 0) return;
    MeritArgs ma = merit_args(ctx, 1, nullptr);
    SFGS_DISPATCH_G(ctx->G_built, auto kern = k_merit<GG>; hipLaunchKernelGGL(kern, dim3(mgrid), dim3(TPB), 0, st, ma,
                    ctx->nslices));
  };
  HIPCHK(hipEventRecord(ctx->ev0, st));
  if (!mixed)
  {
    SweepArgs<double> sa = sweep_args<double>(ctx, par, ctx->W.p, ctx->Wd.p, ctx->B.p, ctx->rho.p, ctx->pfri.p,
                                              ctx->pcoh.p, ctx->shift.p, ctx->U.p, ctx->Um.p, par->diagepsilon);
    if (flow)
    {
      float fms = 0.0f;
      CHK(run_flow(ctx, sa, ctx->R.p, ctx->Rb.p, maxiter, nomerit, par->epsilon, par->meritval, ctx->hist.p, hc, &fms));
      pl.sweep_ms += fms;
      pl.launches += 1;
    }
    else if (bar)
    {
      float bms = 0.0f;
      CHK(run_bar(ctx, sa, ctx->R.p, ctx->Rb.p, maxiter, nomerit, par->epsilon, par->meritval, par->reverse,
                  ctx->hist.p, hc, &bms));
      pl.sweep_ms += bms;
      pl.launches += 1;
    }
    else
      CHK(run_gs<double>(ctx, sa, ctx->R.p, ctx->Rb.p, maxiter, nomerit, par->epsilon, par->meritval, par->reverse,
                         ctx->hist.p, pl, hc));
    if (n > 0 && hc.final_buf == 1)
      HIPCHK(hipMemcpyAsync(ctx->R.p, ctx->Rb.p, sizeof(double)*4*(size_t)n, hipMemcpyDeviceToDevice, st));
    if (n > 0 && hc.final_U == 1) merit_U();
  }
  else
  {
    // FP32 sweeps (FP64 merit, errors and termination), then FP64 sweeps
    CHK(build_f32_mirror(ctx));
    if (n > 0)
    {
      hipLaunchKernelGGL(k_to_f32, dim3(nblk_ll(4LL*n)), dim3(TPB), 0, st, 4LL*n, ctx->R.p, ctx->R32.p);
      hipLaunchKernelGGL(k_to_f32, dim3(nblk_ll(3LL*n)), dim3(TPB), 0, st, 3LL*n, ctx->U.p, ctx->U32.p);
    }
    // FP32 resolves a relative change of about 1e-7, and the FP64 merit of an
    // FP32 fixed point is of order 1e-12 (1e-11 on ellipsoid packings): the
    // local tolerance, the relative error and the merit of the FP32 phase stop
    // at 1e-6, 1e-6 and 1e-10; the FP64 phase takes the solve down to epsilon
    // and meritval
    const double de32 = std::max(par->diagepsilon, 1e-6);
    const double eps32 = std::max(par->epsilon, 1e-6), merit32 = std::max(par->meritval, 1e-10);
    SweepArgs<float> sa = sweep_args<float>(ctx, par, ctx->W32.p, ctx->Wd32.p, ctx->B32.p, ctx->rho32.p,
                                            ctx->pfri32.p, ctx->pcoh32.p, ctx->shift32.p, ctx->U32.p, ctx->Um32.p, de32);
    if (flow)
    {
      float fms = 0.0f;
      CHK(run_flow(ctx, sa, ctx->R32.p, ctx->R32b.p, maxiter32, nomerit, eps32, merit32, ctx->hist.p, hc,
                   &fms));
      pl.sweep_ms += fms;
      pl.launches += 1;
    }
    else if (bar)
    {
      float bms = 0.0f;
      CHK(run_bar(ctx, sa, ctx->R32.p, ctx->R32b.p, maxiter32, nomerit, eps32, merit32, par->reverse,
                  ctx->hist.p, hc, &bms));
      pl.sweep_ms += bms;
      pl.launches += 1;
    }
    else
      CHK(run_gs<float>(ctx, sa, ctx->R32.p, ctx->R32b.p, maxiter32, nomerit, eps32, merit32,
                        par->reverse, ctx->hist.p, pl, hc));
    if (n > 0)
    {
      float *Rf = hc.final_buf == 1 ? ctx->R32b.p : ctx->R32.p;
      hipLaunchKernelGGL(k_to_f64, dim3(nblk_ll(4LL*n)), dim3(TPB), 0, st, 4LL*n, Rf, ctx->R.p);
      if (hc.final_U == 1) merit_U();   // U = B + W R in FP64
      else hipLaunchKernelGGL(k_to_f64, dim3(nblk_ll(3LL*n)), dim3(TPB), 0, st, 3LL*n, ctx->U32.p, ctx->U.p);
    }
    if (refine > 0 && !hc.aborted)
    {
      SweepArgs<double> sd = sweep_args<double>(ctx, par, ctx->W.p, ctx->Wd.p, ctx->B.p, ctx->rho.p, ctx->pfri.p,
                                                ctx->pcoh.p, ctx->shift.p, ctx->U.p, ctx->Um.p, par->diagepsilon);
      // the reference's termination test on FP64 sweeps, not before `refine` of
      // them, within the iterations the FP32 phase left (at least refine)
      CHK(run_gs<double>(ctx, sd, ctx->R.p, ctx->Rb.p, maxiter - hc.iters, nomerit, par->epsilon, par->meritval,
                         par->reverse, ctx->hist.p + 2*(size_t)maxiter, pl, hr, refine));
      if (n > 0 && hr.final_buf == 1)
        HIPCHK(hipMemcpyAsync(ctx->R.p, ctx->Rb.p, sizeof(double)*4*(size_t)n, hipMemcpyDeviceToDevice, st));
      if (n > 0 && hr.final_U == 1) merit_U();
    }
    else if (!nomerit && mgrid > 0)
    {
      // no FP64 sweeps: the FP64 merit of the FP32 state, with the merit's update of U (mrf.c:61-76)
      MeritArgs ma = merit_args(ctx, 1, nullptr);
      SFGS_DISPATCH_G(ctx->G_built, auto kern = k_merit<GG>; hipLaunchKernelGGL(kern, dim3(mgrid), dim3(TPB), 0, st, ma,
                      ctx->nslices));
    }
  }
  HIPCHK(hipEventRecord(ctx->ev1, st));
  unsigned long long fk = 0;
  int fc = 0;
  std::vector<double> hist(4*(size_t)maxiter + 2, 0.0), mp((size_t)std::max(mgrid, 1), 0.0);
  const bool refined = mixed && refine > 0 && !hc.aborted;
  HIPCHK(hipMemcpyAsync(&fk, ctx->failkey.p, sizeof(fk), hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(&fc, ctx->failcount.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(hist.data(), ctx->hist.p, sizeof(double)*hist.size(), hipMemcpyDeviceToHost, st));
  if (mixed && !refined && !nomerit && mgrid > 0)
    HIPCHK(hipMemcpyAsync(mp.data(), ctx->mpartials.p, sizeof(double)*(size_t)mgrid, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipGetLastError());
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));

  // histories of the first phase: errors at hist[k], merits at hist[bound + k]
  const int bound1 = mixed ? maxiter32 : maxiter;
  int iters = hc.iters;
  double rel_error = hc.error, merit = hc.merit;
  std::vector<double> rer((size_t)maxiter, 0.0), mer((size_t)maxiter, 0.0);
  for (int k = 0; k < std::min(hc.iters, maxiter); k ++)
  {
    rer[k] = hist[k];
    mer[k] = hist[(size_t)bound1 + k];
  }
  if (refined)
  {
    // the FP64 phase, appended: its bound is maxiter - hc.iters, so the total stays within maxiter
    const double *h2 = hist.data() + 2*(size_t)maxiter;
    const int bound2 = maxiter - hc.iters;
    for (int k = 0; k < hr.iters && iters + k < maxiter; k ++)
    {
      rer[(size_t)iters + k] = h2[k];
      mer[(size_t)iters + k] = h2[(size_t)bound2 + k];
    }
    iters += hr.iters;
    rel_error = hr.error;
    merit = hr.merit;
  }
  else if (mixed && !nomerit)
  {
    double up = 0.0;
    for (int k = 0; k < mgrid; k ++) up += mp[k];
    double lo = ctx->energy < DBL_EPSILON ? 1.0 : ctx->energy;
    merit = 0.5*up / lo;
  }
  if (rerhist) std::memcpy(rerhist, rer.data(), sizeof(double)*(size_t)std::min(iters, maxiter));
  if (merhist) std::memcpy(merhist, mer.data(), sizeof(double)*(size_t)std::min(iters, maxiter));
  if (rep)
  {
    std::memset(rep, 0, sizeof(*rep));
    rep->iters = iters;
    rep->rel_error = rel_error;
    rep->merit = merit;
    rep->ncolours = ncol;
    rep->fail_rows = fc;
    rep->aborted = hc.aborted || hr.aborted;
    rep->solve_ms = ms;
    rep->sweep_ms = pl.sweep_ms;
    rep->sweep_launches = pl.launches;
    rep->row_updates = (long long)(hc.sweeps + hr.sweeps) * n;   // every sweep run, the merit's extra one included
    rep->nnz = ctx->nnz;
    rep->fp64_sweeps = mixed ? hr.sweeps : hc.sweeps;
    int lps = pl.gc < ncol ? 1 : 0;
    for (int c = 0; c < std::min(pl.gc, ncol); c ++) lps += pl.grid[c] > 0 ? 1 : 0;
    rep->launches_per_sweep = flow || bar ? 0 : lps;
    rep->error = fk ? (int)(fk & 3ull) : SFGS_GS_OK;
    if (!rep->aborted && iters >= maxiter) rep->error = SFGS_GS_DIVERGED;
  }
  return 0;
}

int SFGS_Merit(sfgs_ctx *ctx, int update_U, double *merit, double *conmerit)
{
  if (!ctx) return SFGS_E_ARG;
  if (!ctx->assembled) { ctx->err = "merit before LOCDYN assembly"; return SFGS_E_STATE; }
  if (ctx->diag_only) { ctx->err = "the store holds a penalty solve, not an assembled system"; return SFGS_E_STATE; }
  if (ctx->upmin && update_U)
  {
    ctx->err = "the local dynamics hold the diagonal blocks only (SFGS_Locdyn_Update_Min)";
    return SFGS_E_STATE;
  }
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int n = ctx->nc, mgrid = (ctx->nslices + TPB/SLICE - 1) / (TPB/SLICE);
  CHK(ensure(ctx, ctx->mpartials, (size_t)std::max(mgrid, 1)));
  DBuf<double> cm;
  if (conmerit) { CHK(ensure(ctx, cm, (size_t)n)); }
  MeritArgs ma = merit_args(ctx, update_U, conmerit ? cm.p : nullptr);
  if (mgrid > 0) { SFGS_DISPATCH_G(ctx->G_built, auto kern = k_merit<GG>; hipLaunchKernelGGL(kern, dim3(mgrid), dim3(TPB), 0, st, ma, ctx->nslices)); }
  std::vector<double> mp((size_t)std::max(mgrid, 1), 0.0);
  std::vector<double> hcm(conmerit ? (size_t)n : 0);
  std::vector<int> hperm(conmerit ? (size_t)n : 0);
  if (mgrid > 0) HIPCHK(hipMemcpyAsync(mp.data(), ctx->mpartials.p, sizeof(double)*(size_t)mgrid, hipMemcpyDeviceToHost, st));
  if (conmerit && n > 0)
  {
    HIPCHK(hipMemcpyAsync(hcm.data(), cm.p, sizeof(double)*(size_t)n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(hperm.data(), ctx->perm.p, sizeof(int)*(size_t)n, hipMemcpyDeviceToHost, st));
  }
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipGetLastError());
  release(cm);
  double up = 0.0;
  for (int k = 0; k < mgrid; k ++) up += mp[k];
  double lo = ctx->energy < DBL_EPSILON ? 1.0 : ctx->energy;
  if (merit) *merit = 0.5*up / lo;
  if (conmerit) for (int p = 0; p < n; p ++) conmerit[hperm[p]] = hcm[p] / lo;
  return 0;
}

static int compute_features(sfgs_ctx *ctx, int natural)
{
  const int n = ctx->nc;
  CHK(ensure(ctx, ctx->feat, 18*(size_t)std::max(n, 1)));
  if (n > 0)
  {
    const double *gap = ctx->gap.p;
    hipLaunchKernelGGL(k_features, dim3(nblk(n)), dim3(TPB), 0, ctx->stream, n, ctx->perm.p, gap, ctx->B.p, ctx->Wd.p,
                       ctx->R.p, ctx->rowlen.p, ctx->pfri.p, natural, ctx->feat.p);
  }
  return 0;
}

int SFGS_Features(sfgs_ctx *ctx, float *features)
{
  if (!ctx || !features) return SFGS_E_ARG;
  if (!ctx->assembled || !ctx->from_bodies || ctx->diag_only || ctx->upmin)
  {
    ctx->err = "features need a system assembled from bodies (all blocks)";
    return SFGS_E_STATE;
  }
  HIPCHK(hipSetDevice(ctx->device));
  CHK(compute_features(ctx, 1));
  if (ctx->nc > 0)
    HIPCHK(hipMemcpyAsync(features, ctx->feat.p, sizeof(float)*18*(size_t)ctx->nc, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return 0;
}

int SFGS_Warm_Start(sfgs_ctx *ctx)
{
  if (!ctx) return SFGS_E_ARG;
  if (!ctx->assembled || !ctx->from_bodies || ctx->diag_only || ctx->upmin)
  {
    ctx->err = "warm start needs a system assembled from bodies (all blocks)";
    return SFGS_E_STATE;
  }
  HIPCHK(hipSetDevice(ctx->device));
  CHK(compute_features(ctx, 0));
  return sfgs_mlp_apply(ctx, ctx->feat.p, ctx->nc, ctx->R.p);
}

int SFGS_Set_Schedule(sfgs_ctx *ctx, int schedule)
{
  if (!ctx || schedule < 0 || schedule > 3) return SFGS_E_ARG;
  ctx->schedule = schedule;
  return 0;
}

int SFGS_Set_Sweep_Block(sfgs_ctx *ctx, int waves_per_block)
{
  if (!ctx || (waves_per_block != 0 && waves_per_block != 1 && waves_per_block != 2 && waves_per_block != 4))
    return SFGS_E_ARG;
  ctx->wpb = waves_per_block;
  return 0;
}

int SFGS_Set_Layout(sfgs_ctx *ctx, int lanes_per_row)
{
  if (!ctx || (lanes_per_row != 1 && lanes_per_row != 2 && lanes_per_row != 4 && lanes_per_row != 8)) return SFGS_E_ARG;
  ctx->G = lanes_per_row;
  return 0;
}

int SFGS_Set_Row_Order(sfgs_ctx *ctx, int order)
{
  if (!ctx || order < 0 || order > 3) return SFGS_E_ARG;
  ctx->row_order = order;
  return 0;
}

int SFGS_Get_RU(sfgs_ctx *ctx, double *R, double *U)
{
  if (!ctx) return SFGS_E_ARG;
  if (!ctx->assembled) { ctx->err = "no system"; return SFGS_E_STATE; }
  HIPCHK(hipSetDevice(ctx->device));
  const int n = ctx->nc;
  if (n == 0) return 0;
  CHK(ensure(ctx, ctx->Rn, 3*(size_t)n));
  CHK(ensure(ctx, ctx->Un, 3*(size_t)n));
  hipLaunchKernelGGL(k_unperm, dim3(nblk(n)), dim3(TPB), 0, ctx->stream, n, ctx->perm.p, ctx->R.p, ctx->U.p, ctx->Rn.p, ctx->Un.p);
  if (R) HIPCHK(hipMemcpyAsync(R, ctx->Rn.p, sizeof(double)*3*(size_t)n, hipMemcpyDeviceToHost, ctx->stream));
  if (U) HIPCHK(hipMemcpyAsync(U, ctx->Un.p, sizeof(double)*3*(size_t)n, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return 0;
}

int SFGS_Update_Cohesion(sfgs_ctx *ctx)
{
  if (!ctx) return SFGS_E_ARG;
  if (!ctx->assembled) { ctx->err = "no system"; return SFGS_E_STATE; }
  if (!ctx->from_bodies) { ctx->err = "cohesion states belong to uploaded contacts, not to an uploaded system"; return SFGS_E_STATE; }
  HIPCHK(hipSetDevice(ctx->device));
  const int n = ctx->nc;
  if (n == 0) return 0;
  if (ctx->coh.cap < (size_t)n || ctx->fri.cap < (size_t)n || ctx->perm.cap < (size_t)n || ctx->R.cap < 4*(size_t)n)
  {
    ctx->err = "no contact records for the solved system";
    return SFGS_E_STATE;
  }
  hipLaunchKernelGGL(k_cohesion, dim3(nblk(n)), dim3(TPB), 0, ctx->stream, n, ctx->perm.p, ctx->R.p, ctx->fri.p,
                     ctx->coh.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return 0;
}

int SFGS_Get_Cohesion(sfgs_ctx *ctx, double *cohesion)
{
  if (!ctx || (ctx->nc > 0 && !cohesion)) return SFGS_E_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  const size_t n = (size_t)ctx->nc;
  if (n == 0) return 0;
  if (ctx->coh.cap < n) { ctx->err = "no contact records"; return SFGS_E_STATE; }
  HIPCHK(hipMemcpyAsync(cohesion, ctx->coh.p, sizeof(double)*n, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return 0;
}

int SFGS_Bodyspace_Product(sfgs_ctx *ctx, const double *x, double *z)
{
  if (!ctx || (ctx->nc > 0 && (!x || !z))) return SFGS_E_ARG;
  if (!ctx->assembled || !ctx->from_bodies || ctx->diag_only)
  {
    ctx->err = "body-space products need a LOCDYN assembly from bodies";
    return SFGS_E_STATE;
  }
  HIPCHK(hipSetDevice(ctx->device));
  const int n = ctx->nc;
  if (n == 0) return 0;
  DBuf<double> xn, xp, zn;
  int rc = upload(ctx, xn, x, 3*(size_t)n);
  if (!rc) rc = ensure(ctx, xp, 3*(size_t)n);
  if (!rc) rc = ensure(ctx, zn, 3*(size_t)n);
  if (!rc)
  {
    hipLaunchKernelGGL(k_perm3, dim3(nblk(n)), dim3(TPB), 0, ctx->stream, n, ctx->perm.p, xn.p, xp.p);
    rc = bodyspace_product(ctx, xp.p, 3, 2, nullptr, nullptr, 0.0, 0.0, 0.0, zn.p);
  }
  if (!rc)
  {
    hipError_t e = hipMemcpyAsync(z, zn.p, sizeof(double)*3*(size_t)n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) { ctx->err = hipGetErrorString(e); rc = SFGS_E_HIP; }
  }
  release(xn); release(xp); release(zn);
  return rc;
}

int SFGS_Set_Cohesion(sfgs_ctx *ctx, const double *cohesion)
{
  if (!ctx || (ctx->nc > 0 && !cohesion)) return SFGS_E_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  const size_t n = (size_t)ctx->nc;
  if (n == 0) return 0;
  if (ctx->coh.cap < n) { ctx->err = "no contact records"; return SFGS_E_STATE; }
  HIPCHK(hipMemcpyAsync(ctx->coh.p, cohesion, sizeof(double)*n, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return 0;
}

int SFGS_Get_Sizes(sfgs_ctx *ctx, int *n, long long *nnz, int *ncolours)
{
  if (!ctx) return SFGS_E_ARG;
  if (n) *n = ctx->nc;
  if (nnz) *nnz = ctx->nnz;
  if (ncolours) *ncolours = ctx->ncolours;
  return 0;
}

int SFGS_Get_Locdyn(sfgs_ctx *ctx, int *rowptr, int *col, int *via, double *W, double *Wd,
                    double *A, double *B, double *V, double *rho, double *free_energy, int *colour)
{
  if (!ctx) return SFGS_E_ARG;
  if (!ctx->assembled) { ctx->err = "no system"; return SFGS_E_STATE; }
  if (ctx->diag_only) { ctx->err = "the store holds a penalty solve, not an assembled system"; return SFGS_E_STATE; }
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int n = ctx->nc;
  const long long nnz = ctx->nnz;
  std::vector<int> hperm((size_t)n), hrowptr((size_t)n + 1), hrow0((size_t)ctx->nslices), hnr((size_t)ctx->nslices);
  std::vector<long long> hoff((size_t)ctx->nslices + 1);
  std::vector<double> hWd(9*(size_t)n), hA(9*(size_t)n), hB(3*(size_t)n), hV(3*(size_t)n), hrho((size_t)n);
  std::vector<double> hW(9*(size_t)ctx->sell_blocks);
  std::vector<int> hcol((size_t)nnz), hvia((size_t)nnz), hwcol((size_t)ctx->sell_blocks);
  if (n > 0)
  {
    HIPCHK(hipMemcpyAsync(hperm.data(), ctx->perm.p, sizeof(int)*(size_t)n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(hrowptr.data(), ctx->rowptr.p, sizeof(int)*((size_t)n + 1), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(hWd.data(), ctx->Wd.p, sizeof(double)*9*(size_t)n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(hA.data(), ctx->A.p, sizeof(double)*9*(size_t)n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(hB.data(), ctx->B.p, sizeof(double)*3*(size_t)n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(hV.data(), ctx->V.p, sizeof(double)*3*(size_t)n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(hrho.data(), ctx->rho.p, sizeof(double)*(size_t)n, hipMemcpyDeviceToHost, st));
  }
  if (nnz > 0)
  {
    HIPCHK(hipMemcpyAsync(hcol.data(), ctx->col.p, sizeof(int)*(size_t)nnz, hipMemcpyDeviceToHost, st));
    if (ctx->from_bodies) HIPCHK(hipMemcpyAsync(hvia.data(), ctx->via.p, sizeof(int)*(size_t)nnz, hipMemcpyDeviceToHost, st));
  }
  if (ctx->nslices > 0)
  {
    HIPCHK(hipMemcpyAsync(hrow0.data(), ctx->slice_row0.p, sizeof(int)*(size_t)ctx->nslices, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(hnr.data(), ctx->slice_nrows.p, sizeof(int)*(size_t)ctx->nslices, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(hoff.data(), ctx->slice_off.p, sizeof(long long)*(size_t)ctx->nslices, hipMemcpyDeviceToHost, st));
  }
  if (ctx->sell_blocks > 0)
  {
    HIPCHK(hipMemcpyAsync(hW.data(), ctx->W.p, sizeof(double)*9*(size_t)ctx->sell_blocks, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(hwcol.data(), ctx->wcol.p, sizeof(int)*(size_t)ctx->sell_blocks, hipMemcpyDeviceToHost, st));
  }
  std::vector<int> hcolour((size_t)n);
  if (n > 0) HIPCHK(hipMemcpyAsync(hcolour.data(), ctx->colour.p, sizeof(int)*(size_t)n, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (rowptr) std::memcpy(rowptr, hrowptr.data(), sizeof(int)*((size_t)n + 1));
  if (col) std::memcpy(col, hcol.data(), sizeof(int)*(size_t)nnz);
  if (via) std::memcpy(via, hvia.data(), sizeof(int)*(size_t)nnz);
  if (colour) std::memcpy(colour, hcolour.data(), sizeof(int)*(size_t)n);
  if (free_energy) *free_energy = ctx->energy;
  for (int p = 0; p < n; p ++)
  {
    int i = hperm[p];
    for (int e = 0; e < 9; e ++)
    {
      if (Wd) Wd[9*(size_t)i + e] = hWd[(size_t)e*n + p];
      if (A) A[9*(size_t)i + e] = hA[(size_t)e*n + p];
    }
    for (int k = 0; k < 3; k ++)
    {
      if (B) B[3*(size_t)i + k] = hB[(size_t)k*n + p];
      if (V) V[3*(size_t)i + k] = hV[(size_t)k*n + p];
    }
    if (rho) rho[i] = hrho[p];
  }
  if (W)
  {
    // SELL slot k of row p holds the k-th block of the natural row perm[p]
    for (int s = 0; s < ctx->nslices; s ++)
      for (int l = 0; l < hnr[s]; l ++)
      {
        int p = hrow0[s] + l, i = hperm[p];
        for (int q = hrowptr[i], k = 0; q < hrowptr[i + 1]; q ++, k ++)
        {
          int G = ctx->G_built, m = k / G, lane = l*G + k % G;
          for (int e = 0; e < 9; e ++)
            W[9*(size_t)q + e] = hW[9*(size_t)hoff[s] + (9*(size_t)m + e)*SLICE + lane];
        }
      }
  }
  return 0;
}

int SFGS_Colour(sfgs_ctx *ctx, int n, const int *rowptr, const int *col, int *colour)
{
  if (!ctx || n < 0 || !rowptr || (n > 0 && !colour)) return SFGS_E_ARG;
  long long nnz = rowptr[n];
  if (rowptr[0] != 0 || nnz < 0 || (nnz > 0 && !col)) return SFGS_E_ARG;
  for (long long q = 0; q < nnz; q ++) if (col[q] < 0 || col[q] >= n) return SFGS_E_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  CHK(upload(ctx, ctx->rowptr, rowptr, (size_t)n + 1));
  CHK(upload(ctx, ctx->col, col, (size_t)nnz));
  ctx->assembled = 0;
  CHK(colour_graph(ctx, n));
  if (n > 0) HIPCHK(hipMemcpyAsync(colour, ctx->colour.p, sizeof(int)*(size_t)n, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return ctx->ncolours;
}

int SFGS_Colour_Static(sfgs_ctx *ctx, int n, long long nnz, const int *rowptr, const int *col, int *colour)
{
  if (!ctx || n < 0 || nnz < 0 || nnz > INT_MAX || (n > 0 && (!rowptr || !colour)) || (nnz > 0 && !col))
    return SFGS_E_ARG;
  if (n > SFGS_STATIC_COLOUR_MAX)
  {
    ctx->err = "SFGS_Colour_Static colours at most " + std::to_string(SFGS_STATIC_COLOUR_MAX) + " vertices";
    return SFGS_E_ARG;
  }
  if (!ctx->static_scratch) { ctx->err = "no address for the static scratch array on this device"; return SFGS_E_HIP; }
  if (n == 0) return 0;
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  int *next = ctx->static_scratch, *counter = ctx->static_scratch + SFGS_STATIC_COLOUR_MAX;
  int h[2] = {0, 0};
  // counters cleared by a kernel (no memset node if a caller captures the stream)
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, st, 4, 0, counter);
  hipLaunchKernelGGL(k_check_csr, dim3(nblk(n)), dim3(TPB), 0, st, n, nnz, rowptr, col, counter + 2);
  HIPCHK(hipMemcpyAsync(h, counter + 2, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipGetLastError());
  if (h[0]) { ctx->err = std::to_string(h[0]) + " rows of the graph are not valid CSR"; return SFGS_E_ARG; }
  hipLaunchKernelGGL(k_fill_int, dim3(nblk(n)), dim3(TPB), 0, st, n, -1, colour);
  hipLaunchKernelGGL(k_fill_int, dim3(nblk(n)), dim3(TPB), 0, st, n, -1, next);
  for (int round = 0; ; round ++)
  {
    if (round > 100000) { ctx->err = "colouring did not finish"; return SFGS_E_STATE; }
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, st, 1, 0, counter);
    hipLaunchKernelGGL(k_jp_round, dim3(nblk(n)), dim3(TPB), 0, st, n, rowptr, col, colour, next, counter, counter + 1);
    hipLaunchKernelGGL(k_copy_int, dim3(nblk(n)), dim3(TPB), 0, st, n, next, colour);
    HIPCHK(hipMemcpyAsync(h, counter, 2*sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    HIPCHK(hipGetLastError());
    if (h[0] == 0) break;
  }
  if (h[1] >= SFGS_MAX_COLOURS) { ctx->err = "graph needs more than 256 colours"; return SFGS_E_COLOURS; }
  return h[1] + 1;
}

int SFGS_Body_Update(sfgs_ctx *ctx, double step, double *velo_out, double *force_out)
{
  if (!ctx) return SFGS_E_ARG;
  if (!ctx->assembled || !ctx->from_bodies) { ctx->err = "body update needs a system assembled from bodies"; return SFGS_E_STATE; }
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int nb = ctx->nb;
  if (nb == 0) return 0;
  DBuf<double> &out = ctx->body_out;
  CHK(ensure(ctx, out, 12*(size_t)nb));
  BodyView bv{ctx->kind.p, ctx->rot.p, ctx->x0.p, ctx->jinv.p, ctx->minv.p, ctx->velo.p, ctx->velo0.p};
  ConView cv{ctx->master.p, ctx->slave.p, ctx->mpnt.p, ctx->spnt.p, ctx->base.p, ctx->gap.p, ctx->fri.p,
             ctx->rest.p, ctx->coh.p, ctx->Rn.p, ctx->Un.p};
  hipLaunchKernelGGL(k_body_update, dim3(nblk(nb)), dim3(TPB), 0, st, bv, cv, nb, ctx->bptr.p, ctx->blist.p,
                     ctx->iperm.p, ctx->R.p, step, out.p, out.p + 6*(size_t)nb);
  if (velo_out) HIPCHK(hipMemcpyAsync(velo_out, out.p, sizeof(double)*6*(size_t)nb, hipMemcpyDeviceToHost, st));
  if (force_out) HIPCHK(hipMemcpyAsync(force_out, out.p + 6*(size_t)nb, sizeof(double)*6*(size_t)nb, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipGetLastError());
  return 0;
}


// PENALTY_Solve (pes.c:247-281)
int SFGS_Penalty_Solve(sfgs_ctx *ctx, double step, int implicit, const double *spring, const double *dashpot,
                       const double *hpow)
{
  if (!ctx || !(step > 0.0)) return SFGS_E_ARG;
  if (!ctx->have_bodies) { ctx->err = "no bodies"; return SFGS_E_STATE; }
  const bool given = spring || dashpot || hpow;
  if (given && !(spring && dashpot && hpow)) { ctx->err = "give spring, dashpot and hpow together, or none"; return SFGS_E_ARG; }
  const int n = ctx->nc;
  if (!given && !ctx->dom && n > 0)
  {
    ctx->err = "no material table (SFGS_Dom_Upload) for the springs and dashpots";
    return SFGS_E_STATE;
  }
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  CHK(build_incidence(ctx));
  CHK(ensure(ctx, ctx->perm, (size_t)std::max(n, 1)));
  CHK(ensure(ctx, ctx->iperm, (size_t)std::max(n, 1)));
  CHK(ensure(ctx, ctx->R, 4*(size_t)std::max(n, 1)));
  CHK(ensure(ctx, ctx->U, 3*(size_t)std::max(n, 1)));
  if (given && n > 0)
  {
    std::vector<double> kcp(3*(size_t)n);
    for (int i = 0; i < n; i ++) { kcp[3*(size_t)i] = spring[i]; kcp[3*(size_t)i + 1] = dashpot[i]; kcp[3*(size_t)i + 2] = hpow[i]; }
    CHK(upload(ctx, ctx->pen, kcp.data(), kcp.size()));
  }
  if (!given && n > 0) CHK(upload(ctx, ctx->pen, ctx->defmat, 6));   // the default material's six values
  if (n > 0)
  {
    BodyView bv{ctx->kind.p, ctx->rot.p, ctx->x0.p, ctx->jinv.p, ctx->minv.p, ctx->velo.p, ctx->velo0.p};
    ConView cv{ctx->master.p, ctx->slave.p, ctx->mpnt.p, ctx->spnt.p, ctx->base.p, ctx->gap.p, ctx->fri.p,
               ctx->rest.p, ctx->coh.p, ctx->Rn.p, ctx->Un.p};
    PenMat pm{given ? ctx->pen.p : nullptr, ctx->surf.p, ctx->mat_s.p, ctx->mat_v.p, given ? nullptr : ctx->pen.p,
              ctx->nmat};
    hipLaunchKernelGGL(k_identity, dim3(nblk(n)), dim3(TPB), 0, st, n, ctx->perm.p, ctx->iperm.p);
    hipLaunchKernelGGL(k_penalty, dim3(nblk(n)), dim3(TPB), 0, st, bv, cv, pm, n, implicit ? 1 : 0, step, ctx->R.p,
                       ctx->U.p, ctx->coh.p);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  ctx->step = step;
  ctx->assembled = 1;
  ctx->from_bodies = 1;
  ctx->diag_only = 1;
  ctx->upmin = 0;
  ctx->nnz = 0;
  ctx->ncolours = 0;
  ctx->mirror_valid = 0;
  return 0;
}

} // extern "C"

int sfgs::reassemble(sfgs_ctx *ctx)
{
  if (!ctx->from_bodies) { ctx->err = "re-assembly needs a system assembled from bodies"; return SFGS_E_STATE; }
  return SFGS_Locdyn_Update_Begin(ctx, ctx->step, ctx->dynamic);
}
