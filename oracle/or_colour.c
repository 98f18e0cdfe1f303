/* or_colour.c -- contact-graph colouring used to order the Gauss-Seidel sweep.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * No reference counterpart: the reference sweeps its contact list serially
 * (bgs.c:950) and colours only MPI ranks (greedy first-fit, bgs.c:90-104).
 * The GPU schedule colours contacts with Jones-Plassmann rounds under the
 * priority key (row length << 32 | OR_Hash32(index)), ties broken by the larger
 * index, each selected vertex taking the smallest colour absent from its
 * already coloured neighbours.  That result equals sequential greedy first-fit
 * over vertices in descending priority, which is what is computed here; the
 * GPU colouring must match it bit for bit.
 *
 * Parity: pinned in tests/test_oracle_colour.py -- proper colouring (no edge
 * inside a colour class), first-fit minimality (every vertex of colour c has a
 * higher-priority neighbour of each colour below c), a brute-force Python
 * greedy on small graphs, and the hash against its published test values. */
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* 32-bit integer mixer "lowbias32" (C. Wellons, hash-prospector) */
unsigned OR_Hash32(unsigned x)
{
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

static const int *g_rowptr;
static int cmp_prio(const void *pa, const void *pb)
{
  int a = *(const int*)pa, b = *(const int*)pb;
  unsigned long long ka = ((unsigned long long)(unsigned)(g_rowptr[a + 1] - g_rowptr[a]) << 32) | OR_Hash32((unsigned)a);
  unsigned long long kb = ((unsigned long long)(unsigned)(g_rowptr[b + 1] - g_rowptr[b]) << 32) | OR_Hash32((unsigned)b);
  if (ka != kb) return ka > kb ? -1 : 1;
  return a > b ? -1 : (a < b ? 1 : 0);
}

int OR_Colour(int n, const int *rowptr, const int *col, int *colour)
{
  int *ord = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1)), ncol = 0, maxdeg = 0;
  if (!ord) return -1;
  for (int i = 0; i < n; i ++)
  {
    ord[i] = i;
    colour[i] = -1;
    if (rowptr[i + 1] - rowptr[i] > maxdeg) maxdeg = rowptr[i + 1] - rowptr[i];
  }
  g_rowptr = rowptr;
  qsort(ord, (size_t)n, sizeof(int), cmp_prio);
  unsigned char *used = calloc((size_t)maxdeg + 2, 1);
  if (!used) { free(ord); return -1; }
  for (int k = 0; k < n; k ++)
  {
    int v = ord[k], c = 0;
    for (int p = rowptr[v]; p < rowptr[v + 1]; p ++)
    {
      int cu = colour[col[p]];
      if (cu >= 0 && cu <= maxdeg) used[cu] = 1;
    }
    while (used[c]) c ++;
    colour[v] = c;
    if (c + 1 > ncol) ncol = c + 1;
    for (int p = rowptr[v]; p < rowptr[v + 1]; p ++)
    {
      int cu = colour[col[p]];
      if (cu >= 0 && cu <= maxdeg) used[cu] = 0;
    }
  }
  free(used);
  free(ord);
  return ncol;
}

void OR_Colour_Order(int n, const int *colour, int ncolours, int *order, int *colptr)
{
  memset(colptr, 0, sizeof(int) * (size_t)(ncolours + 1));
  for (int i = 0; i < n; i ++) colptr[colour[i] + 1] ++;
  for (int c = 0; c < ncolours; c ++) colptr[c + 1] += colptr[c];
  int *fill = malloc(sizeof(int) * (size_t)(ncolours > 0 ? ncolours : 1));
  memcpy(fill, colptr, sizeof(int) * (size_t)ncolours);
  for (int i = 0; i < n; i ++) order[fill[colour[i]] ++] = i;
  free(fill);
}
