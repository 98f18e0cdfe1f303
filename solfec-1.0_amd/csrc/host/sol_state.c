/* sol_state.c -- state file of the domain for restart and warm-start persistence.
 *
 * Records follow the reference's state writers field by field:
 *   bodies      BODY_Write_State (bod.c:2202-2221, IOVER >= 3): kind, configuration
 *               size, dofs, configuration (R 9, x 3), velocity (W 3, v 3), energy
 *               (4 values for RIG, none for OBS; kept as zeros here);
 *   constraints write_constraint (dio.c:19-59, IOVER 4, kind CONTACT): id, kind, R,
 *               U, V, point, base, merit, master and slave ids, the surface material
 *               state (SURFACE_MATERIAL_Write_State, sps.c:286-294: material index,
 *               state size -- 1 when the material is cohesive, sps.c:65-69 -- and
 *               the contact's cohesion state), area, gap, surface pair.
 * Encoding: XDR (RFC 4506), which the reference's non-HDF5 PBF files use: big-endian,
 * 4-byte ints (a short takes 4 bytes), 8-byte IEEE doubles, strings as length +
 * bytes padded to 4.  The container is this library's own: a header (the string
 * "SOLFEC-1.0 STATE", IOVER, time, step, body count, contact count) followed by the
 * body records in body order and the contact records in contact order.  Ids are the
 * body index + 1 (0 = no slave), the contact index + 1.  The reference's default
 * build wraps the same records in HDF5 groups (pbf.c:36-700), which this build does
 * not write. */
#define _GNU_SOURCE
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "solfec.h"
#include "sol_internal.h"
#include "sol_geom.h"

#define STATE_MAGIC "SOLFEC-1.0 STATE"
#define STATE_IOVER 4
#define REF_OBS 0   /* bod.h:156 */
#define REF_RIG 1

static void put_u32(FILE *f, unsigned v)
{
  unsigned char b[4] = {(unsigned char)(v >> 24), (unsigned char)(v >> 16), (unsigned char)(v >> 8), (unsigned char)v};
  fwrite(b, 1, 4, f);
}
static void put_f64(FILE *f, double d)
{
  unsigned long long v;
  memcpy(&v, &d, 8);
  put_u32(f, (unsigned)(v >> 32));
  put_u32(f, (unsigned)v);
}
static void put_f64s(FILE *f, const double *d, int n) { for (int k = 0; k < n; k ++) put_f64(f, d[k]); }
static void put_str(FILE *f, const char *s)
{
  unsigned n = (unsigned)strlen(s), pad = (4 - n % 4) % 4;
  put_u32(f, n);
  fwrite(s, 1, n, f);
  for (unsigned k = 0; k < pad; k ++) fputc(0, f);
}

typedef struct { const unsigned char *p; size_t n, pos; int bad; } RD;
static unsigned get_u32(RD *r)
{
  if (r->pos + 4 > r->n) { r->bad = 1; return 0; }
  const unsigned char *b = r->p + r->pos;
  r->pos += 4;
  return ((unsigned)b[0] << 24) | ((unsigned)b[1] << 16) | ((unsigned)b[2] << 8) | (unsigned)b[3];
}
static int get_i32(RD *r) { return (int)get_u32(r); }
static double get_f64(RD *r)
{
  unsigned long long hi = get_u32(r), lo = get_u32(r), v = (hi << 32) | lo;
  double d;
  memcpy(&d, &v, 8);
  return d;
}
static void get_f64s(RD *r, double *d, int n) { for (int k = 0; k < n; k ++) d[k] = get_f64(r); }

/* index of a contact's surface pair in the material table, the default last */
static int material_index(SOLFEC *sol, int s1, int s2)
{
  int a = s1 < s2 ? s1 : s2, b = s1 < s2 ? s2 : s1;
  for (int k = 0; k < sol->nmat; k ++) if (sol->mat[k].s1 == a && sol->mat[k].s2 == b) return k;
  return sol->nmat;
}

int SOLFEC_Write_State(SOLFEC *sol, const char *path)
{
  if (!sol || !path) return SOL_ERR_ARG;
  int rc = sol_dev_pull(sol);
  if (rc) return rc;
  const int nb = sol->nb, nc = sol->nc;
  /* V and the per-contact merit of the last assembled step, when it is this contact list */
  double *V = calloc(3*(size_t)nc + 3, sizeof(double)), *merit = calloc((size_t)nc + 1, sizeof(double));
  if (!V || !merit) { free(V); free(merit); return SOL_ERR_ARG; }
  int n = -1;
  if (sol->gpu && sol->locdyn_current && !SFGS_Get_Sizes(sol->gpu, &n, NULL, NULL) && n == nc && nc > 0)
  {
    double m;
    if (SFGS_Get_Locdyn(sol->gpu, NULL, NULL, NULL, NULL, NULL, NULL, NULL, V, NULL, NULL, NULL) ||
        SFGS_Merit(sol->gpu, 0, &m, merit))
    {
      memset(V, 0, sizeof(double)*3*(size_t)nc);
      memset(merit, 0, sizeof(double)*(size_t)nc);
    }
  }
  FILE *f = fopen(path, "wb");
  if (!f)
  {
    free(V); free(merit);
    snprintf(sol->err, sizeof(sol->err), "cannot open %s for writing", path);
    return SOL_ERR_ARG;
  }
  put_str(f, STATE_MAGIC);
  put_u32(f, STATE_IOVER);
  put_f64(f, sol->time);
  put_f64(f, sol->step);
  put_u32(f, (unsigned)nb);
  put_u32(f, (unsigned)nc);
  static const double zero4[4] = {0.0, 0.0, 0.0, 0.0};
  for (int b = 0; b < nb; b ++)
  {
    int obs = sol->kind[b] == SOL_OBS;
    put_u32(f, obs ? REF_OBS : REF_RIG);
    put_u32(f, 12);
    put_u32(f, 6);
    put_f64s(f, sol->rot + 9*(size_t)b, 9);
    put_f64s(f, sol->x + 3*(size_t)b, 3);
    put_f64s(f, sol->velo + 6*(size_t)b, 6);
    if (!obs) put_f64s(f, zero4, 4);
  }
  for (int i = 0; i < nc; i ++)
  {
    int m = sol->c_master[i], s = sol->c_slave[i];
    put_u32(f, (unsigned)(i + 1));
    put_u32(f, 0);                                 /* CONTACT */
    put_f64s(f, sol->c_R + 3*(size_t)i, 3);
    put_f64s(f, sol->c_U + 3*(size_t)i, 3);
    put_f64s(f, V + 3*(size_t)i, 3);
    put_f64s(f, sol->c_point + 3*(size_t)i, 3);
    put_f64s(f, sol->c_base + 9*(size_t)i, 9);
    put_f64(f, merit[i]);
    put_u32(f, (unsigned)(m + 1));
    put_u32(f, s >= 0 ? (unsigned)(s + 1) : 0u);
    const int k = material_index(sol, sol->surf[m], s >= 0 ? sol->surf[s] : sol->surf[m]);
    const MATPAIR *mp = k < sol->nmat ? &sol->mat[k] : &sol->defmat;
    put_u32(f, (unsigned)k);
    put_u32(f, mp->cohesion > 0.0 ? 1u : 0u);     /* material state size */
    if (mp->cohesion > 0.0) put_f64(f, sol->c_coh[i]);
    put_f64(f, 1.0);                               /* area (goc.c:450) */
    put_f64(f, sol->c_gap[i]);
    put_u32(f, (unsigned)sol->surf[m]);
    put_u32(f, (unsigned)(s >= 0 ? sol->surf[s] : sol->surf[m]));
  }
  int bad = ferror(f);
  bad |= fclose(f);
  free(V); free(merit);
  if (bad) { snprintf(sol->err, sizeof(sol->err), "write error on %s", path); return SOL_ERR_ARG; }
  return SOL_OK;
}

int SOLFEC_Read_State(SOLFEC *sol, const char *path)
{
  if (!sol || !path) return SOL_ERR_ARG;
  FILE *f = fopen(path, "rb");
  if (!f) { snprintf(sol->err, sizeof(sol->err), "cannot open %s", path); return SOL_ERR_ARG; }
  fseek(f, 0, SEEK_END);
  long len = ftell(f);
  fseek(f, 0, SEEK_SET);
  unsigned char *buf = malloc(len > 0 ? (size_t)len : 1);
  if (!buf || (len > 0 && fread(buf, 1, (size_t)len, f) != (size_t)len)) { free(buf); fclose(f); return SOL_ERR_ARG; }
  fclose(f);
  RD r = {buf, (size_t)(len > 0 ? len : 0), 0, 0};
  int rc = SOL_ERR_ARG;
  unsigned slen = get_u32(&r);
  if (r.bad || slen != strlen(STATE_MAGIC) || r.pos + 4*((slen + 3)/4) > r.n ||
      memcmp(buf + r.pos, STATE_MAGIC, slen))
  {
    snprintf(sol->err, sizeof(sol->err), "%s is not a state file of this library", path);
    free(buf);
    return SOL_ERR_ARG;
  }
  r.pos += 4*((slen + 3)/4);
  unsigned iover = get_u32(&r);
  double time = get_f64(&r), step = get_f64(&r);
  int nb = get_i32(&r), nc = get_i32(&r);
  /* a contact record takes at least 224 bytes, a body record 84 */
  if (r.bad || iover != STATE_IOVER || nb != sol->nb || nc < 0 || step != sol->step ||
      (size_t)nc > (r.n - r.pos) / 224 || (size_t)nb > (r.n - r.pos) / 84)
  {
    snprintf(sol->err, sizeof(sol->err), "state file %s does not match this domain (version %u, %d bodies, step %g)",
             path, iover, nb, step);
    free(buf);
    return SOL_ERR_ARG;
  }
  sol_dev_edit(sol);
  /* decode into temporaries first, so that a bad file leaves the domain as it was */
  double *body = malloc(sizeof(double)*18*(size_t)(nb > 0 ? nb : 1));
  int *cm = malloc(sizeof(int)*(size_t)(nc + 1)), *cs = malloc(sizeof(int)*(size_t)(nc + 1));
  double *con = malloc(sizeof(double)*20*(size_t)(nc + 1));
  if (!body || !cm || !cs || !con) goto done;
  for (int b = 0; b < nb; b ++)
  {
    int kind = get_i32(&r), conf = get_i32(&r), dofs = get_i32(&r);
    int obs = sol->kind[b] == SOL_OBS;
    if (r.bad || kind != (obs ? REF_OBS : REF_RIG) || conf != 12 || dofs != 6)
    {
      snprintf(sol->err, sizeof(sol->err), "body %d of %s does not match this domain", b, path);
      goto done;
    }
    get_f64s(&r, body + 18*(size_t)b, 18);
    if (!obs) { double e[4]; get_f64s(&r, e, 4); }
  }
  for (int i = 0; i < nc; i ++)
  {
    double *c = con + 20*(size_t)i;
    unsigned id = get_u32(&r);
    int kind = get_i32(&r);
    get_f64s(&r, c, 3);          /* R */
    get_f64s(&r, c + 3, 3);      /* U */
    double V[3], merit;
    get_f64s(&r, V, 3);
    get_f64s(&r, c + 6, 3);      /* point */
    get_f64s(&r, c + 9, 9);      /* base */
    merit = get_f64(&r);
    (void)merit;
    unsigned mid = get_u32(&r), sid = get_u32(&r);
    (void)get_u32(&r);           /* material index: the table is the domain's */
    int ssize = get_i32(&r);
    c[19] = 0.0;                 /* cohesion state */
    if (ssize > 0 && ssize < 64) { double st[64]; get_f64s(&r, st, ssize); c[19] = st[0]; }
    (void)get_f64(&r);           /* area */
    c[18] = get_f64(&r);         /* gap */
    (void)get_u32(&r); (void)get_u32(&r);   /* surface pair */
    if (r.bad || id != (unsigned)(i + 1) || kind != 0 || mid < 1 || mid > (unsigned)nb || sid < 1 ||
        sid > (unsigned)nb || mid == sid || ssize < 0 || ssize >= 64)
    {
      snprintf(sol->err, sizeof(sol->err), "contact %d of %s is not valid here", i, path);
      goto done;
    }
    cm[i] = (int)mid - 1;
    cs[i] = (int)sid - 1;
  }
  if (r.pos != r.n) { snprintf(sol->err, sizeof(sol->err), "%s has trailing data", path); goto done; }
  if (sol_ensure_contacts(sol, nc > 0 ? nc : 1)) { rc = SOL_ERR_MEM; goto done; }
  /* install */
  sol->time = time;
  for (int b = 0; b < nb; b ++)
  {
    const double *q = body + 18*(size_t)b;
    memcpy(sol->rot + 9*(size_t)b, q, 9*sizeof(double));
    memcpy(sol->x + 3*(size_t)b, q + 9, 3*sizeof(double));
    memcpy(sol->velo + 6*(size_t)b, q + 12, 6*sizeof(double));
  }
  sol->nc = nc;
  for (int i = 0; i < nc; i ++)
  {
    const double *c = con + 20*(size_t)i;
    int m = cm[i], s = cs[i];
    sol->c_master[i] = m;
    sol->c_slave[i] = s;
    memcpy(sol->c_R + 3*(size_t)i, c, 3*sizeof(double));
    memcpy(sol->c_U + 3*(size_t)i, c + 3, 3*sizeof(double));
    memcpy(sol->c_point + 3*(size_t)i, c + 6, 3*sizeof(double));
    memcpy(sol->c_base + 9*(size_t)i, c + 9, 9*sizeof(double));
    sol->c_gap[i] = c[18];
    /* referential points from the stored point; every contact is re-evaluated at the next step begin */
    geom_ref_point(sol->rot + 9*(size_t)m, sol->x + 3*(size_t)m, sol->x0 + 3*(size_t)m, c + 6, sol->c_mpnt + 3*(size_t)i);
    geom_ref_point(sol->rot + 9*(size_t)s, sol->x + 3*(size_t)s, sol->x0 + 3*(size_t)s, c + 6, sol->c_spnt + 3*(size_t)i);
    int k = material_index(sol, sol->surf[m], sol->surf[s]);
    const MATPAIR *mp = k < sol->nmat ? &sol->mat[k] : &sol->defmat;
    sol->c_fri[i] = mp->friction;
    sol->c_rest[i] = mp->restitution;
    /* cohesive when the material is and the state holds cohesion (sps.c:72-80) */
    sol->c_coh[i] = mp->cohesion > 0.0 && c[19] > 0.0 ? c[19] : 0.0;
  }
  rc = SOL_OK;
done:
  free(buf); free(body); free(cm); free(cs); free(con);
  return rc;
}
