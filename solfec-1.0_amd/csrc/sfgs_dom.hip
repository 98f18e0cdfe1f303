// sfgs_dom.hip -- device-resident time stepping of rigid spheres, ellipsoids
// and fixed planes: DOM_Update_Begin / DOM_Update_End (dom.c:3630, 3818) on
// the GPU, so that a run loop keeps bodies and contacts in HBM.
//
// Step begin (one call):
//   k_dom_begin     rigid half step of every body (bod.c:1313-1373);
//   k_con_update    geometry of every existing contact (update_contact,
//                   dom.c:430-483), flag kept / removed;
//   k_con_compact   kept contacts in their old order, R and U carried;
//   k_hash_insert   existing body pairs into an open-addressing set;
//   k_bounds_*      bounding box of the convex bodies;
//   k_cell_*        uniform grid of cell 2 r_max + 2 eps, bodies per cell;
//   k_cand_count / k_cand_write
//                   per body a, the bodies b > a of the 27 neighbour cells
//                   whose bounding spheres overlap and whose pair is not a
//                   contact yet, in ascending order of b;
//   k_narrow        orientation (the lexicographically larger centre is the
//                   master, goc.c:421-455) and the convex-pair test;
//   k_plane_flags   plane x convex pairs, plane-major;
//   k_append_*      new contacts after the kept ones, R = U = 0 (mem.c:77,86).
// Step end: R, U back to the contact list, the velocity update by the
// reactions (k_body_update) and k_dom_end (bod.c:1425-1492).
//
// The geometry and the integration come from sol_geom.h, which the host layer
// compiles too; this file is built without floating-point contraction, so the
// device loop computes the host loop's bits.  Work per step: O(bodies +
// contacts + candidates), all in flat arrays; the few counts that size the
// next arrays are read back (one small copy each).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <string>
#include <utility>
#include <vector>
#include "sfgs_host.hpp"
#include "sfgs_prims.hpp"
#include "sol_geom.h"

using namespace sfgs;

namespace {

constexpr unsigned long long EMPTY = ~0ULL;

// bodies as the geometry needs them
struct DomView {
  const int *kind, *shape, *surf;
  const double *rad, *plane, *sca, *rot0, *rot, *x, *x0;
};

struct MatView {
  int nmat;
  const int *s;        // 2 per material, smaller surface first
  const double *v;     // friction, restitution, cohesion, spring, dashpot, hpow
  double def[3];
  int first;           // the domain time is 0: new contacts take the material's cohesion
};

// geometry record of an existing contact: spnt 3, mpnt 3, normal 3, gap, then
// its cohesion state; a cohesive contact carries its previous referential
// points in the first six entries instead (it keeps them, dom.c:452-458)
constexpr int GEO = 11;

// contact record outputs of set_contact
struct ConOut {
  int *master, *slave;
  double *R, *U, *point, *mpnt, *spnt, *base, *gap, *fri, *rest, *coh;
};

__device__ int convex_pair(const DomView &d, int a, int b, double *sp, double *mp, double *nrm, double *g)
{
  if (d.shape[a] == 0 && d.shape[b] == 0)
    return geom_sphere_pair(d.x + 3*(size_t)a, d.rad[a], d.x + 3*(size_t)b, d.rad[b], sp, mp, nrm, g);
  double Qa[9], Sa[3], Qb[9], Sb[3];
  geom_frame(d.shape[a], d.rot + 9*(size_t)a, d.rot0 + 9*(size_t)a, d.sca + 3*(size_t)a, d.rad[a], Qa, Sa);
  geom_frame(d.shape[b], d.rot + 9*(size_t)b, d.rot0 + 9*(size_t)b, d.sca + 3*(size_t)b, d.rad[b], Qb, Sb);
  return geom_ellip_pair(d.x + 3*(size_t)a, Qa, Sa, d.x + 3*(size_t)b, Qb, Sb, sp, mp, nrm, g);
}

__device__ int plane_convex(const DomView &d, int pl, int s, double *sp, double *mp, double *nrm, double *g)
{
  if (d.shape[s] == 2)
  {
    double Q[9], S[3];
    geom_frame(2, d.rot + 9*(size_t)s, d.rot0 + 9*(size_t)s, d.sca + 3*(size_t)s, d.rad[s], Q, S);
    return geom_plane_ellip(d.plane + 6*(size_t)pl, d.x + 3*(size_t)s, Q, S, sp, mp, nrm, g);
  }
  return geom_plane_sphere(d.plane + 6*(size_t)pl, d.x + 3*(size_t)s, d.rad[s], sp, mp, nrm, g);
}

// geometry of a contact given as (slave, master); G = spnt 3, mpnt 3, normal 3, gap
__device__ int contact_geometry(const DomView &d, int s, int m, double *G)
{
  if (s < 0) return 0;   // no two-body pair: not a contact of this loop
  if (d.shape[s] == 1) return plane_convex(d, s, m, G, G + 3, G + 6, G + 9);
  return convex_pair(d, s, m, G, G + 3, G + 6, G + 9);
}

__device__ const double* material_of(const DomView &d, const MatView &mv, int master, int slave)
{
  int s1 = d.surf[master], s2 = d.surf[slave], a = s1 < s2 ? s1 : s2, b = s1 < s2 ? s2 : s1;
  const double *v = mv.def;
  for (int k = 0; k < mv.nmat; k ++)
    if (mv.s[2*k] == a && mv.s[2*k + 1] == b) { v = mv.v + 6*k; break; }
  return v;
}

// set_contact of the host layer: points, referential points, frame, gap,
// material; a new contact (fresh) takes the cohesion state of
// SURFACE_MATERIAL_Transfer (sps.c:264-282), an existing one keeps coh
__device__ void put_contact(const DomView &d, const MatView &mv, const ConOut &o, int i, int master, int slave,
                            const double *mp, const double *sp, const double *nrm, double g, bool fresh, double coh)
{
  o.master[i] = master;
  o.slave[i] = slave;
  for (int k = 0; k < 3; k ++) o.point[3*(size_t)i + k] = mp[k];
  geom_ref_point(d.rot + 9*(size_t)master, d.x + 3*(size_t)master, d.x0 + 3*(size_t)master, mp, o.mpnt + 3*(size_t)i);
  geom_ref_point(d.rot + 9*(size_t)slave, d.x + 3*(size_t)slave, d.x0 + 3*(size_t)slave, sp, o.spnt + 3*(size_t)i);
  geom_localbase(nrm, o.base + 9*(size_t)i);
  o.gap[i] = g;
  const double *v = material_of(d, mv, master, slave);
  o.fri[i] = v[0];
  o.rest[i] = v[1];
  o.coh[i] = fresh ? (mv.first && v[2] > 0.0 ? v[2] : 0.0) : coh;   // area = 1 for these contacts (goc.c:450)
}

// set_cohesive of the host layer: the referential points Mp, Sp are kept and the
// spatial point is the master's original point in the current configuration
__device__ void put_cohesive(const DomView &d, const MatView &mv, const ConOut &o, int i, int master, int slave,
                             const double *Mp, const double *Sp, const double *nrm, double g, double coh)
{
  o.master[i] = master;
  o.slave[i] = slave;
  geom_cur_point(d.rot + 9*(size_t)master, d.x + 3*(size_t)master, d.x0 + 3*(size_t)master, Mp, o.point + 3*(size_t)i);
  for (int k = 0; k < 3; k ++) { o.mpnt[3*(size_t)i + k] = Mp[k]; o.spnt[3*(size_t)i + k] = Sp[k]; }
  geom_localbase(nrm, o.base + 9*(size_t)i);
  o.gap[i] = g;
  const double *v = material_of(d, mv, master, slave);
  o.fri[i] = v[0];
  o.rest[i] = v[1];
  o.coh[i] = coh;
}

__device__ bool pair_known(const unsigned long long *keys, unsigned long long mask, int a, int b)
{
  unsigned long long key = geom_pair_key(a, b), h = geom_mix64(key) & mask;
  while (keys[h] != EMPTY)
  {
    if (keys[h] == key) return true;
    h = (h + 1) & mask;
  }
  return false;
}

// dynamic = 0: the quasi-static step (bod.c:1588-1614), velo0 left alone
__global__ void k_dom_begin(int nb, const int *kind, double h, double g0, double g1, double g2, const double *mass,
                            const double *minv, const double *J, const double *I, double *velo, double *velo0,
                            double *rot, double *x, double *aux, int dynamic)
{
  int b = blockIdx.x*blockDim.x + threadIdx.x;
  if (b >= nb) return;
  if (kind[b] == 1)
  {
    if (dynamic) for (int k = 0; k < 6; k ++) velo0[6*(size_t)b + k] = velo[6*(size_t)b + k];
    return;
  }
  const double grav[3] = {g0, g1, g2};
  if (!dynamic)
  {
    geom_rig_static_begin(h, grav, mass[b], minv[b], J + 9*(size_t)b, I + 9*(size_t)b, velo + 6*(size_t)b);
    return;
  }
  geom_rig_step_begin(h, grav, mass[b], minv[b], J + 9*(size_t)b, I + 9*(size_t)b, velo + 6*(size_t)b,
                      velo0 + 6*(size_t)b, rot + 9*(size_t)b, x + 3*(size_t)b, aux + 3*(size_t)b);
}

__global__ void k_dom_end(int nb, const int *kind, double h, const double *I, const double *bout, double *velo,
                          double *force, double *rot, double *x, double *aux, int dynamic)
{
  int b = blockIdx.x*blockDim.x + threadIdx.x;
  if (b >= nb) return;
  for (int k = 0; k < 6; k ++)
  {
    velo[6*(size_t)b + k] = bout[6*(size_t)b + k];
    force[6*(size_t)b + k] = bout[6*(size_t)nb + 6*(size_t)b + k];
  }
  if (kind[b] == 1) return;
  if (!dynamic)
  {
    geom_rig_static_end(h, velo + 6*(size_t)b, rot + 9*(size_t)b, x + 3*(size_t)b);
    return;
  }
  geom_rig_step_end(h, I + 9*(size_t)b, force + 6*(size_t)b, velo + 6*(size_t)b, rot + 9*(size_t)b,
                    x + 3*(size_t)b, aux + 3*(size_t)b);
}

// geometry of the existing contacts; a cohesive one is kept when separated,
// with its previous normal and gap (dom.c:452-458)
__global__ void k_con_update(int nc, DomView d, const int *master, const int *slave, const double *coh,
                             const double *mpnt, const double *spnt, const double *base, const double *gap,
                             double *geo, int *ok)
{
  int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i > nc) return;
  if (i == nc) { ok[nc] = 0; return; }   // scan sentinel
  double *G = geo + GEO*(size_t)i;
  const int in = contact_geometry(d, slave[i], master[i], G);
  const double c = coh[i];
  G[10] = c;
  if (c > 0.0)
  {
    for (int k = 0; k < 3; k ++) { G[k] = spnt[3*(size_t)i + k]; G[3 + k] = mpnt[3*(size_t)i + k]; }
    if (!in)
    {
      for (int k = 0; k < 3; k ++) G[6 + k] = base[9*(size_t)i + 6 + k];
      G[9] = gap[i];
    }
    ok[i] = 1;
  }
  else ok[i] = in;
}

__global__ void k_con_compact(int nc, DomView d, MatView mv, ConOut o, const int *ok, const int *pos,
                              const int *master, const int *slave, const double *R, const double *U, const double *geo)
{
  int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i >= nc || !ok[i]) return;
  int k = pos[i];
  for (int c = 0; c < 3; c ++) { o.R[3*(size_t)k + c] = R[3*(size_t)i + c]; o.U[3*(size_t)k + c] = U[3*(size_t)i + c]; }
  const double *G = geo + GEO*(size_t)i;
  if (G[10] > 0.0) put_cohesive(d, mv, o, k, master[i], slave[i], G + 3, G, G + 6, G[9], G[10]);
  else put_contact(d, mv, o, k, master[i], slave[i], G + 3, G, G + 6, G[9], false, 0.0);
}

__global__ void k_hash_clear(size_t cap, unsigned long long *keys)
{
  size_t i = (size_t)blockIdx.x*blockDim.x + threadIdx.x;
  if (i < cap) keys[i] = EMPTY;
}

__global__ void k_hash_insert(int nc, const int *ok, const int *master, const int *slave, unsigned long long *keys,
                              unsigned long long mask)
{
  int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i >= nc || !ok[i]) return;
  unsigned long long key = geom_pair_key(master[i], slave[i]), h = geom_mix64(key) & mask;
  for (;;)
  {
    unsigned long long prev = atomicCAS(&keys[h], EMPTY, key);
    if (prev == EMPTY || prev == key) return;
    h = (h + 1) & mask;
  }
}

// bounding box of the convex bodies and the largest radius: per block, then one block
__global__ void k_bounds_part(int nb, const int *shape, const double *x, double *part)
{
  __shared__ double s[6][TPB];
  double v[6] = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
  for (int b = blockIdx.x*blockDim.x + threadIdx.x; b < nb; b += gridDim.x*blockDim.x)
    if (shape[b] != 1)
      for (int k = 0; k < 3; k ++)
      {
        double c = x[3*(size_t)b + k];
        v[k] = fmin(v[k], c);
        v[3 + k] = fmax(v[3 + k], c);
      }
  for (int k = 0; k < 6; k ++) s[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int w = TPB/2; w > 0; w >>= 1)
  {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 6; k ++)
        s[k][threadIdx.x] = k < 3 ? fmin(s[k][threadIdx.x], s[k][threadIdx.x + w])
                                  : fmax(s[k][threadIdx.x], s[k][threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 6; k ++) part[6*(size_t)blockIdx.x + k] = s[k][0];
}

// one block of TPB threads
__global__ void k_bounds_final(int nparts, const double *part, double *out)
{
  __shared__ double s[6][TPB];
  double v[6] = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
  for (int p = threadIdx.x; p < nparts; p += TPB)
    for (int k = 0; k < 6; k ++)
      v[k] = k < 3 ? fmin(v[k], part[6*(size_t)p + k]) : fmax(v[k], part[6*(size_t)p + k]);
  for (int k = 0; k < 6; k ++) s[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int w = TPB/2; w > 0; w >>= 1)
  {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 6; k ++)
        s[k][threadIdx.x] = k < 3 ? fmin(s[k][threadIdx.x], s[k][threadIdx.x + w])
                                  : fmax(s[k][threadIdx.x], s[k][threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 6; k ++) out[k] = s[k][0];
}

struct Grid {
  double lo[3], cell;
  int dim[3];
};

__device__ int cell_of(const Grid &g, const double *c, int *ci)
{
  for (int k = 0; k < 3; k ++) ci[k] = geom_cell_coord((c[k] - g.lo[k])/g.cell, g.dim[k]);
  return (ci[2]*g.dim[1] + ci[1])*g.dim[0] + ci[0];
}

__global__ void k_cell_count(int nb, const int *shape, const double *x, Grid g, int *cnt)
{
  int b = blockIdx.x*blockDim.x + threadIdx.x;
  if (b >= nb || shape[b] == 1) return;
  int ci[3];
  atomicAdd(&cnt[cell_of(g, x + 3*(size_t)b, ci)], 1);
}

__global__ void k_cell_fill(int nb, const int *shape, const double *x, Grid g, const int *start, int *cur, int *body)
{
  int b = blockIdx.x*blockDim.x + threadIdx.x;
  if (b >= nb || shape[b] == 1) return;
  int ci[3], c = cell_of(g, x + 3*(size_t)b, ci);
  body[start[c] + atomicAdd(&cur[c], 1)] = b;
}

// the new-pair candidates of body a; out == nullptr only counts them, else at
// most cap are written
__device__ int cand_visit(int a, const DomView &d, const Grid &g, const int *start, const int *body,
                          const unsigned long long *keys, unsigned long long mask, int *out, int cap = 0)
{
  int ci[3], k = 0;
  const double *xa = d.x + 3*(size_t)a;
  cell_of(g, xa, ci);
  for (int dz = -1; dz <= 1; dz ++)
    for (int dy = -1; dy <= 1; dy ++)
      for (int dx = -1; dx <= 1; dx ++)
      {
        int cx = ci[0] + dx, cy = ci[1] + dy, cz = ci[2] + dz;
        if (cx < 0 || cy < 0 || cz < 0 || cx >= g.dim[0] || cy >= g.dim[1] || cz >= g.dim[2]) continue;
        int c = (cz*g.dim[1] + cy)*g.dim[0] + cx;
        for (int q = start[c]; q < start[c + 1]; q ++)
        {
          int b = body[q];
          if (b <= a) continue;
          const double *xb = d.x + 3*(size_t)b;
          double dx_ = xb[0] - xa[0], dy_ = xb[1] - xa[1], dz_ = xb[2] - xa[2],
                 e_ = d.rad[a] + d.rad[b] + SOL_GEOMETRIC_EPSILON;
          if (dx_*dx_ + dy_*dy_ + dz_*dz_ >= e_*e_) continue;
          if (pair_known(keys, mask, a, b)) continue;
          if (out && k < cap) out[k] = b;
          k ++;
        }
      }
  return k;
}

__global__ void k_cand_count(int nb, DomView d, Grid g, const int *start, const int *body,
                             const unsigned long long *keys, unsigned long long mask, int *cnt)
{
  int a = blockIdx.x*blockDim.x + threadIdx.x;
  if (a > nb) return;
  if (a == nb) { cnt[nb] = 0; return; }
  cnt[a] = d.shape[a] == 1 ? 0 : cand_visit(a, d, g, start, body, keys, mask, nullptr);
}

__global__ void k_cand_write(int nb, DomView d, Grid g, const int *start, const int *body,
                             const unsigned long long *keys, unsigned long long mask, const int *off,
                             int *cand_a, int *cand_b)
{
  int a = blockIdx.x*blockDim.x + threadIdx.x;
  if (a >= nb || d.shape[a] == 1) return;
  int o = off[a], n = off[a + 1] - o;
  if (n == 0) return;
  int *cb = cand_b + o;
  cand_visit(a, d, g, start, body, keys, mask, cb, n);
  for (int q = 1; q < n; q ++)   // ascending b
  {
    int v = cb[q], r = q;
    while (r > 0 && cb[r - 1] > v) { cb[r] = cb[r - 1]; r --; }
    cb[r] = v;
  }
  for (int q = 0; q < n; q ++) cand_a[o + q] = a;
}

__global__ void k_narrow(int m, DomView d, const int *cand_a, const int *cand_b, int *ok, int *cm, int *cs, double *geo)
{
  int q = blockIdx.x*blockDim.x + threadIdx.x;
  if (q > m) return;
  if (q == m) { ok[m] = 0; return; }
  int a = cand_a[q], b = cand_b[q], slave = a, master = b;
  if (geom_pntcmp(d.x + 3*(size_t)a, d.x + 3*(size_t)b) > 0) { slave = b; master = a; }
  double *G = geo + 10*(size_t)q;
  ok[q] = convex_pair(d, slave, master, G, G + 3, G + 6, G + 9);
  cm[q] = master;
  cs[q] = slave;
}

__global__ void k_plane_flags(int np, int nb, DomView d, const int *planes, const unsigned long long *keys,
                              unsigned long long mask, int *ok)
{
  long long idx = (long long)blockIdx.x*blockDim.x + threadIdx.x, tot = (long long)np*nb;
  if (idx > tot) return;
  if (idx == tot) { ok[tot] = 0; return; }
  int p = planes[idx / nb], s = (int)(idx % nb);
  ok[idx] = 0;
  if (d.shape[s] == 1) return;
  const double *P = d.plane + 6*(size_t)p, *c = d.x + 3*(size_t)s;
  double dist = (c[0] - P[0])*P[3] + (c[1] - P[1])*P[4] + (c[2] - P[2])*P[5];
  if (dist >= d.rad[s] + SOL_GEOMETRIC_EPSILON) return;
  if (pair_known(keys, mask, p, s)) return;
  double G[10];
  ok[idx] = plane_convex(d, p, s, G, G + 3, G + 6, G + 9);
}

__global__ void k_append_pairs(int m, int base, DomView d, MatView mv, ConOut o, const int *ok, const int *pos,
                               const int *cm, const int *cs, const double *geo)
{
  int q = blockIdx.x*blockDim.x + threadIdx.x;
  if (q >= m || !ok[q]) return;
  int k = base + pos[q];
  for (int c = 0; c < 3; c ++) { o.R[3*(size_t)k + c] = 0.0; o.U[3*(size_t)k + c] = 0.0; }
  const double *G = geo + 10*(size_t)q;
  put_contact(d, mv, o, k, cm[q], cs[q], G + 3, G, G + 6, G[9], true, 0.0);
}

__global__ void k_append_planes(int np, int nb, int base, DomView d, MatView mv, ConOut o, const int *planes,
                                const int *ok, const int *pos)
{
  long long idx = (long long)blockIdx.x*blockDim.x + threadIdx.x;
  if (idx >= (long long)np*nb || !ok[idx]) return;
  int p = planes[idx / nb], s = (int)(idx % nb), k = base + pos[idx];
  double G[10];
  plane_convex(d, p, s, G, G + 3, G + 6, G + 9);
  for (int c = 0; c < 3; c ++) { o.R[3*(size_t)k + c] = 0.0; o.U[3*(size_t)k + c] = 0.0; }
  put_contact(d, mv, o, k, s, p, G + 3, G, G + 6, G[9], true, 0.0);
}

int scan(sfgs_ctx *ctx, const int *in, int *out, int n)
{
  size_t bytes = 0;
  HIPCHK(sfgs::scan_int(nullptr, bytes, in, out, n, ctx->stream));
  CHK(ensure(ctx, ctx->scratch, bytes));
  HIPCHK(sfgs::scan_int(ctx->scratch.p, bytes, in, out, n, ctx->stream));
  return 0;
}

// total of an exclusive scan of n + 1 entries whose last input is 0
int scan_total(sfgs_ctx *ctx, const int *in, int *out, int n, int *total)
{
  CHK(scan(ctx, in, out, n + 1));
  HIPCHK(hipMemcpyAsync(total, out + n, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return 0;
}

DomView dom_view(sfgs_ctx *ctx)
{
  return DomView{ctx->kind.p, ctx->shape.p, ctx->surf.p, ctx->rad.p, ctx->plane.p, ctx->sca.p, ctx->rot0.p,
                 ctx->rot.p, ctx->bx.p, ctx->x0.p};
}

MatView mat_view(sfgs_ctx *ctx)
{
  MatView m;
  m.nmat = ctx->nmat; m.s = ctx->mat_s.p; m.v = ctx->mat_v.p;
  for (int k = 0; k < 3; k ++) m.def[k] = ctx->defmat[k];
  m.first = ctx->dom_time == 0.0;
  return m;
}

} // namespace

void sfgs_dom_release(sfgs_ctx *ctx)
{
  release(ctx->shape); release(ctx->surf); release(ctx->planes);
  release(ctx->rad); release(ctx->plane); release(ctx->sca); release(ctx->rot0); release(ctx->bx);
  release(ctx->J); release(ctx->mass); release(ctx->aux); release(ctx->force); release(ctx->point);
  release(ctx->mat_s); release(ctx->mat_v);
  release(ctx->geo); release(ctx->geo2); release(ctx->dpart);
  release(ctx->okf); release(ctx->okpos); release(ctx->ok2); release(ctx->okpos2); release(ctx->ok3); release(ctx->okpos3);
  release(ctx->cand_m); release(ctx->cand_s); release(ctx->t_master); release(ctx->t_slave);
  release(ctx->t_R); release(ctx->t_U); release(ctx->hkeys);
  release(ctx->cell_cnt); release(ctx->cell_start); release(ctx->cell_body); release(ctx->cell_cur);
  release(ctx->ncand); release(ctx->candoff); release(ctx->cand_a); release(ctx->cand_b);
}

extern "C" {

int SFGS_Dom_Upload(sfgs_ctx *ctx, const SFGS_SHAPES *sh, int nmat, const SFGS_MATERIAL *mat,
                    const SFGS_MATERIAL *defmat, const double *point)
{
  if (!ctx || !sh || nmat < 0 || (nmat > 0 && !mat) || !defmat) return SFGS_E_ARG;
  if (!ctx->have_bodies) { ctx->err = "bodies must be uploaded before their shapes"; return SFGS_E_STATE; }
  const int nb = ctx->nb, nc = ctx->nc;
  if (nb > 0 && (!sh->shape || !sh->surf || !sh->rad || !sh->plane || !sh->sca || !sh->rot0 || !sh->x || !sh->J ||
                 !sh->mass || !sh->aux)) return SFGS_E_ARG;
  if (nc > 0 && !point) return SFGS_E_ARG;
  std::vector<int> planes;
  double rmax = 0.0;
  for (int b = 0; b < nb; b ++)
  {
    if (sh->shape[b] < 0 || sh->shape[b] > 2)
    {
      ctx->err = "body " + std::to_string(b) + " has an unknown shape";
      return SFGS_E_ARG;
    }
    if (sh->shape[b] == 1) planes.push_back(b);
    else rmax = std::max(rmax, sh->rad[b]);
  }
  if ((long long)planes.size()*nb > (1LL << 30)) { ctx->err = "too many planes"; return SFGS_E_ARG; }
  HIPCHK(hipSetDevice(ctx->device));
  CHK(upload(ctx, ctx->shape, sh->shape, (size_t)nb));
  CHK(upload(ctx, ctx->surf, sh->surf, (size_t)nb));
  CHK(upload(ctx, ctx->rad, sh->rad, (size_t)nb));
  CHK(upload(ctx, ctx->plane, sh->plane, 6*(size_t)nb));
  CHK(upload(ctx, ctx->sca, sh->sca, 3*(size_t)nb));
  CHK(upload(ctx, ctx->rot0, sh->rot0, 9*(size_t)nb));
  CHK(upload(ctx, ctx->bx, sh->x, 3*(size_t)nb));
  CHK(upload(ctx, ctx->J, sh->J, 9*(size_t)nb));
  CHK(upload(ctx, ctx->mass, sh->mass, (size_t)nb));
  CHK(upload(ctx, ctx->aux, sh->aux, 3*(size_t)nb));
  CHK(ensure(ctx, ctx->force, 6*(size_t)nb));
  HIPCHK(hipMemsetAsync(ctx->force.p, 0, sizeof(double)*6*(size_t)nb, ctx->stream));
  CHK(upload(ctx, ctx->planes, planes.data(), planes.size()));
  CHK(upload(ctx, ctx->point, point, 3*(size_t)nc));
  std::vector<int> ms(2*(size_t)nmat);
  std::vector<double> mvv(6*(size_t)nmat);
  for (int k = 0; k < nmat; k ++)
  {
    ms[2*k] = std::min(mat[k].s1, mat[k].s2);
    ms[2*k + 1] = std::max(mat[k].s1, mat[k].s2);
    mvv[6*k] = mat[k].friction; mvv[6*k + 1] = mat[k].restitution; mvv[6*k + 2] = mat[k].cohesion;
    mvv[6*k + 3] = mat[k].spring; mvv[6*k + 4] = mat[k].dashpot; mvv[6*k + 5] = mat[k].hpow;
  }
  CHK(upload(ctx, ctx->mat_s, ms.data(), ms.size()));
  CHK(upload(ctx, ctx->mat_v, mvv.data(), mvv.size()));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->nmat = nmat;
  ctx->defmat[0] = defmat->friction; ctx->defmat[1] = defmat->restitution; ctx->defmat[2] = defmat->cohesion;
  ctx->defmat[3] = defmat->spring; ctx->defmat[4] = defmat->dashpot; ctx->defmat[5] = defmat->hpow;
  ctx->nplanes = (int)planes.size();
  ctx->rmax = rmax;
  ctx->dom = 1;
  return 0;
}

int SFGS_Dom_Update_Begin(sfgs_ctx *ctx, double step, const double *gravity, int *ncon)
{
  if (!ctx || !(step > 0.0) || !gravity) return SFGS_E_ARG;
  if (!ctx->dom) { ctx->err = "no shapes for the current bodies (SFGS_Dom_Upload)"; return SFGS_E_STATE; }
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int nb = ctx->nb, nc = ctx->nc, np = ctx->nplanes;
  const DomView d = dom_view(ctx);
  const MatView mv = mat_view(ctx);

  // rigid half step (velo0 <- velo; R, x to t + h/2; free velocity)
  if (nb > 0)
    hipLaunchKernelGGL(k_dom_begin, dim3(nblk(nb)), dim3(TPB), 0, st, nb, ctx->kind.p, step, gravity[0], gravity[1],
                       gravity[2], ctx->mass.p, ctx->minv.p, ctx->J.p, ctx->jinv.p, ctx->velo.p, ctx->velo0.p,
                       ctx->rot.p, ctx->bx.p, ctx->aux.p, ctx->dom_dynamic);

  // existing contacts: geometry, kept flags and their order-preserving positions
  CHK(ensure(ctx, ctx->geo, GEO*(size_t)nc));
  CHK(ensure(ctx, ctx->okf, (size_t)nc + 1));
  CHK(ensure(ctx, ctx->okpos, (size_t)nc + 1));
  hipLaunchKernelGGL(k_con_update, dim3(nblk(nc + 1)), dim3(TPB), 0, st, nc, d, ctx->master.p, ctx->slave.p,
                     ctx->coh.p, ctx->mpnt.p, ctx->spnt.p, ctx->base.p, ctx->gap.p, ctx->geo.p, ctx->okf.p);
  int kept = 0;
  CHK(scan_total(ctx, ctx->okf.p, ctx->okpos.p, nc, &kept));

  // the kept body pairs, for the broad phase to skip
  size_t cap = 16;
  while (cap < 2*(size_t)kept + 16) cap *= 2;
  CHK(ensure(ctx, ctx->hkeys, cap));
  const unsigned long long mask = cap - 1;
  hipLaunchKernelGGL(k_hash_clear, dim3(nblk((long long)cap)), dim3(TPB), 0, st, cap, ctx->hkeys.p);
  if (nc > 0)
    hipLaunchKernelGGL(k_hash_insert, dim3(nblk(nc)), dim3(TPB), 0, st, nc, ctx->okf.p, ctx->master.p, ctx->slave.p,
                       ctx->hkeys.p, mask);

  // broad phase over the convex bodies
  int m = 0;
  if (ctx->rmax > 0.0)
  {
    const int nparts = std::min(nblk(nb), 1024);
    CHK(ensure(ctx, ctx->dpart, 6*(size_t)nparts + 6));
    hipLaunchKernelGGL(k_bounds_part, dim3(nparts), dim3(TPB), 0, st, nb, ctx->shape.p, ctx->bx.p, ctx->dpart.p);
    hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(TPB), 0, st, nparts, ctx->dpart.p, ctx->dpart.p + 6*(size_t)nparts);
    double bb[6];
    HIPCHK(hipMemcpyAsync(bb, ctx->dpart.p + 6*(size_t)nparts, sizeof(bb), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    Grid g;
    g.cell = 2.0*ctx->rmax + 2.0*SOL_GEOMETRIC_EPSILON;
    long long ncell = 1;
    for (int k = 0; k < 3; k ++) g.lo[k] = bb[k];
    // at most 4 cells per body (a coarser grid for very sparse scenes), sized
    // without integer overflow however far a body has gone, so that every
    // cell index of k_cell_* and k_cand_* fits an int and the cell arrays
    if (geom_grid_dims(bb, bb + 3, 4LL*nb + 1024, &g.cell, g.dim, &ncell))
    {
      ctx->err = "body positions are not finite";
      return SFGS_E_ARG;
    }
    CHK(ensure(ctx, ctx->cell_cnt, (size_t)ncell + 1));
    CHK(ensure(ctx, ctx->cell_start, (size_t)ncell + 1));
    CHK(ensure(ctx, ctx->cell_cur, (size_t)ncell));
    CHK(ensure(ctx, ctx->cell_body, (size_t)nb));
    HIPCHK(hipMemsetAsync(ctx->cell_cnt.p, 0, sizeof(int)*((size_t)ncell + 1), st));
    HIPCHK(hipMemsetAsync(ctx->cell_cur.p, 0, sizeof(int)*(size_t)ncell, st));
    hipLaunchKernelGGL(k_cell_count, dim3(nblk(nb)), dim3(TPB), 0, st, nb, ctx->shape.p, ctx->bx.p, g, ctx->cell_cnt.p);
    CHK(scan(ctx, ctx->cell_cnt.p, ctx->cell_start.p, (int)ncell + 1));
    hipLaunchKernelGGL(k_cell_fill, dim3(nblk(nb)), dim3(TPB), 0, st, nb, ctx->shape.p, ctx->bx.p, g,
                       ctx->cell_start.p, ctx->cell_cur.p, ctx->cell_body.p);
    CHK(ensure(ctx, ctx->ncand, (size_t)nb + 1));
    CHK(ensure(ctx, ctx->candoff, (size_t)nb + 1));
    hipLaunchKernelGGL(k_cand_count, dim3(nblk(nb + 1)), dim3(TPB), 0, st, nb, d, g, ctx->cell_start.p,
                       ctx->cell_body.p, ctx->hkeys.p, mask, ctx->ncand.p);
    CHK(scan_total(ctx, ctx->ncand.p, ctx->candoff.p, nb, &m));
    CHK(ensure(ctx, ctx->cand_a, (size_t)m));
    CHK(ensure(ctx, ctx->cand_b, (size_t)m));
    if (m > 0)
      hipLaunchKernelGGL(k_cand_write, dim3(nblk(nb)), dim3(TPB), 0, st, nb, d, g, ctx->cell_start.p,
                         ctx->cell_body.p, ctx->hkeys.p, mask, ctx->candoff.p, ctx->cand_a.p, ctx->cand_b.p);
  }
  // narrow phase of the candidates
  CHK(ensure(ctx, ctx->ok2, (size_t)m + 1));
  CHK(ensure(ctx, ctx->okpos2, (size_t)m + 1));
  CHK(ensure(ctx, ctx->cand_m, (size_t)m));
  CHK(ensure(ctx, ctx->cand_s, (size_t)m));
  CHK(ensure(ctx, ctx->geo2, 10*(size_t)m));
  hipLaunchKernelGGL(k_narrow, dim3(nblk(m + 1)), dim3(TPB), 0, st, m, d, ctx->cand_a.p, ctx->cand_b.p, ctx->ok2.p,
                     ctx->cand_m.p, ctx->cand_s.p, ctx->geo2.p);
  int n2 = 0;
  CHK(scan_total(ctx, ctx->ok2.p, ctx->okpos2.p, m, &n2));

  // planes against the convex bodies
  const long long npl = (long long)np*nb;
  CHK(ensure(ctx, ctx->ok3, (size_t)npl + 1));
  CHK(ensure(ctx, ctx->okpos3, (size_t)npl + 1));
  hipLaunchKernelGGL(k_plane_flags, dim3(nblk(npl + 1)), dim3(TPB), 0, st, np, nb, d, ctx->planes.p, ctx->hkeys.p,
                     mask, ctx->ok3.p);
  int n3 = 0;
  CHK(scan_total(ctx, ctx->ok3.p, ctx->okpos3.p, (int)npl, &n3));

  // the new contact list: kept, new convex pairs, new plane pairs
  const long long ntot = (long long)kept + n2 + n3;
  if (ntot > (1LL << 30)) { ctx->err = "too many contacts"; return SFGS_E_ARG; }
  const size_t nn = (size_t)ntot;
  CHK(ensure(ctx, ctx->t_master, nn)); CHK(ensure(ctx, ctx->t_slave, nn));
  CHK(ensure(ctx, ctx->t_R, 3*nn)); CHK(ensure(ctx, ctx->t_U, 3*nn));
  // the point and frame arrays are rewritten whole; what the kept contacts need
  // of their old contents is in geo (k_con_update)
  CHK(ensure(ctx, ctx->point, 3*nn)); CHK(ensure(ctx, ctx->mpnt, 3*nn)); CHK(ensure(ctx, ctx->spnt, 3*nn));
  CHK(ensure(ctx, ctx->base, 9*nn)); CHK(ensure(ctx, ctx->gap, nn)); CHK(ensure(ctx, ctx->fri, nn));
  CHK(ensure(ctx, ctx->rest, nn)); CHK(ensure(ctx, ctx->coh, nn));
  ConOut o{ctx->t_master.p, ctx->t_slave.p, ctx->t_R.p, ctx->t_U.p, ctx->point.p, ctx->mpnt.p, ctx->spnt.p,
           ctx->base.p, ctx->gap.p, ctx->fri.p, ctx->rest.p, ctx->coh.p};
  if (nc > 0)
    hipLaunchKernelGGL(k_con_compact, dim3(nblk(nc)), dim3(TPB), 0, st, nc, d, mv, o, ctx->okf.p, ctx->okpos.p,
                       ctx->master.p, ctx->slave.p, ctx->Rn.p, ctx->Un.p, ctx->geo.p);
  if (m > 0)
    hipLaunchKernelGGL(k_append_pairs, dim3(nblk(m)), dim3(TPB), 0, st, m, kept, d, mv, o, ctx->ok2.p, ctx->okpos2.p,
                       ctx->cand_m.p, ctx->cand_s.p, ctx->geo2.p);
  if (npl > 0)
    hipLaunchKernelGGL(k_append_planes, dim3(nblk(npl)), dim3(TPB), 0, st, np, nb, kept + n2, d, mv, o, ctx->planes.p,
                       ctx->ok3.p, ctx->okpos3.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  std::swap(ctx->master, ctx->t_master);
  std::swap(ctx->slave, ctx->t_slave);
  std::swap(ctx->Rn, ctx->t_R);
  std::swap(ctx->Un, ctx->t_U);
  ctx->nc = (int)ntot;
  ctx->assembled = 0;
  if (ncon) *ncon = (int)ntot;
  return 0;
}

int SFGS_Dom_Set_Time(sfgs_ctx *ctx, double time)
{
  if (!ctx || !(time >= 0.0)) return SFGS_E_ARG;
  ctx->dom_time = time;
  return 0;
}

int SFGS_Dom_Set_Analysis(sfgs_ctx *ctx, int dynamic)
{
  if (!ctx || (dynamic != 0 && dynamic != 1)) return SFGS_E_ARG;
  ctx->dom_dynamic = dynamic;
  return 0;
}

int SFGS_Dom_Update_End(sfgs_ctx *ctx, double step)
{
  if (!ctx || !(step > 0.0)) return SFGS_E_ARG;
  if (!ctx->dom) { ctx->err = "no shapes for the current bodies (SFGS_Dom_Upload)"; return SFGS_E_STATE; }
  if (!ctx->assembled) { ctx->err = "step end needs the solved system of the step"; return SFGS_E_STATE; }
  CHK(SFGS_Get_RU(ctx, nullptr, nullptr));
  CHK(SFGS_Body_Update(ctx, step, nullptr, nullptr));
  const int nb = ctx->nb;
  if (nb > 0)
  {
    hipLaunchKernelGGL(k_dom_end, dim3(nblk(nb)), dim3(TPB), 0, ctx->stream, nb, ctx->kind.p, step, ctx->jinv.p,
                       ctx->body_out.p, ctx->velo.p, ctx->force.p, ctx->rot.p, ctx->bx.p, ctx->aux.p,
                       ctx->dom_dynamic);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
  }
  return 0;
}

int SFGS_Dom_Get_Bodies(sfgs_ctx *ctx, double *rot, double *x, double *velo, double *velo0, double *aux, double *force)
{
  if (!ctx) return SFGS_E_ARG;
  if (!ctx->dom) { ctx->err = "no device loop state"; return SFGS_E_STATE; }
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const size_t nb = (size_t)ctx->nb;
  if (nb > 0)
  {
    if (rot) HIPCHK(hipMemcpyAsync(rot, ctx->rot.p, sizeof(double)*9*nb, hipMemcpyDeviceToHost, st));
    if (x) HIPCHK(hipMemcpyAsync(x, ctx->bx.p, sizeof(double)*3*nb, hipMemcpyDeviceToHost, st));
    if (velo) HIPCHK(hipMemcpyAsync(velo, ctx->velo.p, sizeof(double)*6*nb, hipMemcpyDeviceToHost, st));
    if (velo0) HIPCHK(hipMemcpyAsync(velo0, ctx->velo0.p, sizeof(double)*6*nb, hipMemcpyDeviceToHost, st));
    if (aux) HIPCHK(hipMemcpyAsync(aux, ctx->aux.p, sizeof(double)*3*nb, hipMemcpyDeviceToHost, st));
    if (force) HIPCHK(hipMemcpyAsync(force, ctx->force.p, sizeof(double)*6*nb, hipMemcpyDeviceToHost, st));
  }
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

int SFGS_Dom_Get_Contacts(sfgs_ctx *ctx, int *master, int *slave, double *point, double *mpnt, double *spnt,
                          double *base, double *gap, double *friction, double *restitution, double *cohesion,
                          double *R, double *U)
{
  if (!ctx) return SFGS_E_ARG;
  if (!ctx->dom) { ctx->err = "no device loop state"; return SFGS_E_STATE; }
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const size_t n = (size_t)ctx->nc;
  if (n > 0)
  {
    if (master) HIPCHK(hipMemcpyAsync(master, ctx->master.p, sizeof(int)*n, hipMemcpyDeviceToHost, st));
    if (slave) HIPCHK(hipMemcpyAsync(slave, ctx->slave.p, sizeof(int)*n, hipMemcpyDeviceToHost, st));
    if (point) HIPCHK(hipMemcpyAsync(point, ctx->point.p, sizeof(double)*3*n, hipMemcpyDeviceToHost, st));
    if (mpnt) HIPCHK(hipMemcpyAsync(mpnt, ctx->mpnt.p, sizeof(double)*3*n, hipMemcpyDeviceToHost, st));
    if (spnt) HIPCHK(hipMemcpyAsync(spnt, ctx->spnt.p, sizeof(double)*3*n, hipMemcpyDeviceToHost, st));
    if (base) HIPCHK(hipMemcpyAsync(base, ctx->base.p, sizeof(double)*9*n, hipMemcpyDeviceToHost, st));
    if (gap) HIPCHK(hipMemcpyAsync(gap, ctx->gap.p, sizeof(double)*n, hipMemcpyDeviceToHost, st));
    if (friction) HIPCHK(hipMemcpyAsync(friction, ctx->fri.p, sizeof(double)*n, hipMemcpyDeviceToHost, st));
    if (restitution) HIPCHK(hipMemcpyAsync(restitution, ctx->rest.p, sizeof(double)*n, hipMemcpyDeviceToHost, st));
    if (cohesion) HIPCHK(hipMemcpyAsync(cohesion, ctx->coh.p, sizeof(double)*n, hipMemcpyDeviceToHost, st));
    if (R) HIPCHK(hipMemcpyAsync(R, ctx->Rn.p, sizeof(double)*3*n, hipMemcpyDeviceToHost, st));
    if (U) HIPCHK(hipMemcpyAsync(U, ctx->Un.p, sizeof(double)*3*n, hipMemcpyDeviceToHost, st));
  }
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

} // extern "C"
