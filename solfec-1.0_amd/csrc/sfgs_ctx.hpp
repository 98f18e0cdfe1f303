// sfgs_ctx.hpp -- device-resident LOCDYN store (the W/B/R store) of one context.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include <cstdint>
#include "../../include/sfgs.h"

namespace sfgs {

// grow-only device buffer
template <typename T>
struct DBuf {
  T *p = nullptr;
  size_t cap = 0;
  size_t n = 0;
};

// device control block of one solve
struct Ctrl {
  int iters;        // completed Gauss-Seidel iterations (gs->iters)
  int sweeps;       // sweep kernels run so far
  int done;         // terminated (converged, maxiter or abort)
  int aborted;      // FAILURE_EXIT raised in a sweep
  int final_buf;    // R buffer (0 / 1) holding the final reactions
  int final_U;      // 0: U, 1: Um holds the final relative velocities
  int maxiter, nomerit;
  int miniter;      // no convergence stop before this many iterations (the FP64 phase of MIXED)
  double error;     // last relative error
  double merit;     // last merit
  double prev_error;
  double eps, meritval;
  double energy;    // free energy (merit denominator source)
};

struct RowData;  // forward

} // namespace sfgs

struct sfgs_ctx;
void sfgs_mlp_release(sfgs_ctx *ctx);
void sfgs_dom_release(sfgs_ctx *ctx);
void sfgs_newton_release(sfgs_ctx *ctx);
void sfgs_bar_release(sfgs_ctx *ctx);
int sfgs_mlp_apply(sfgs_ctx *ctx, const float *feat, int n, double *R);

struct sfgs_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int *static_scratch = nullptr;   // this device's copy of the file-scope scratch array (SFGS_Colour_Static)

  // ---- bodies, natural order
  int nb = 0;
  sfgs::DBuf<int> kind;
  sfgs::DBuf<double> rot, x0, jinv, minv, velo, velo0;

  // ---- contacts, natural order
  int nc = 0;
  sfgs::DBuf<int> master, slave;
  sfgs::DBuf<double> mpnt, spnt, base, gap, fri, rest, coh, Rn, Un;

  // ---- body -> contact incidence
  sfgs::DBuf<int> bptr, blist, bkeys, bkeys2, bvals, bvals2;

  // ---- natural adjacency CSR
  long long nnz = 0;
  sfgs::DBuf<int> rowptr, rowlen_n, col, via;

  // ---- colouring and permutation
  int ncolours = 0;
  sfgs::DBuf<int> colour, colour2, perm, iperm, counter;
  sfgs::DBuf<unsigned long long> skeys, skeys2;
  sfgs::DBuf<int> svals, svals2;
  std::vector<int> h_colptr, h_cslice;   // rows / slices per colour (host)
  sfgs::DBuf<unsigned char> scratch;     // hipcub temporary storage

  // ---- slices (SELL-64, colour-major)
  int nslices = 0;
  sfgs::DBuf<int> slice_row0, slice_nrows, slice_len;
  sfgs::DBuf<long long> slice_off;       // block offset of slice (in blocks, multiple of 64)
  long long sell_blocks = 0;             // total block slots incl. padding

  // ---- permuted row data (colour-major order)
  sfgs::DBuf<int> rowlen;                // per permuted row
  sfgs::DBuf<double> W, Wd, A, B, V, rho, pfri, pcoh, shift, R, U, Rb, Um;
  int G = 2;                             // lanes per row of the sweep layout (1, 2, 4, 8)
  int wpb = 0;                           // wavefronts per workgroup of the per-colour sweep launches (1, 2, 4;
                                         // 0: by precision, 2 for FP64 solves and 4 for MIXED)
  int wpb_run = 4;                       // the value of the current solve
  int schedule = 0;                      // sweeps: 0 one launch per colour (XCD-aware), 1 persistent dataflow,
                                         // 2 one launch per colour (hardware order), 3 persistent with barriers
  int row_order = 3;                     // inside a colour: 0 by decreasing row length, 1 by contact index,
                                         // 2 by master body block then decreasing row length, 3 by the
                                         // Z-order of the master body's position
  sfgs::DBuf<unsigned long long> bbox;   // row order 3: bounding box of the bodies (order keys)
  sfgs::DBuf<unsigned char> flow_mem;    // dataflow schedule: third R buffer, U pairs, versions, partials
  int G_built = 1;                       // layout of the current store
  sfgs::DBuf<unsigned char> sep;
  sfgs::DBuf<int> wcol;
  sfgs::DBuf<float> W32, Wd32, B32, R32, R32b, U32, Um32, rho32, pfri32, pcoh32, shift32;
  sfgs::DBuf<float> rec32;               // FP32 row records of the FP32 sweeps (five float4 per row)
  int mirror_valid = 0;                  // FP32 copies match the FP64 store
  double energy = 0.0, step = 0.0;
  int dynamic = 1;
  int have_bodies = 0, have_system = 0, assembled = 0, from_bodies = 0;
  int diag_only = 0;                     // the store holds a penalty solve (R, U per contact), no system
  int upmin = 0;                         // diagonal blocks only (SFGS_Locdyn_Update_Min): W products through the bodies
  sfgs::DBuf<double> pen;                // penalty solve: spring, dashpot, hpow per contact, or the default six
  sfgs::DBuf<double> nt_mem;             // Newton solve: linearisation, Krylov vectors, partials
  sfgs::DBuf<unsigned char> bar_mem;     // barrier schedule: colour tables, partials, counter

  // ---- reductions / control
  sfgs::DBuf<double> partials, mpartials, hist;
  sfgs::DBuf<unsigned long long> failkey;
  sfgs::DBuf<int> failcount;
  sfgs::DBuf<sfgs::Ctrl> ctrl;
  sfgs::DBuf<int> fin;                    // ticket counter of the launch that takes a sweep's termination test
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
  std::vector<hipEvent_t> prof_ev;        // per-launch sweep timing (profile mode)
  sfgs::DBuf<double> body_out;            // body update results (velo, force)
  sfgs::DBuf<double> bsv;                 // body-space products: h M^-1 H' x per body (6 per body)
  sfgs::DBuf<float> feat;                 // learned warm start: per-row features (18 per row)
  void *mlp = nullptr;                    // learned warm start model (sfgs_mlp.hip)

  // ---- independent scenes (batched solve, sfgs_batch.hip).  With nscenes > 1
  // rows are ordered by (scene, colour, -row length, i) and h_colptr / h_cslice
  // hold rows / slices per group g = scene*ncolours + colour.
  int nscenes = 1;
  sfgs::DBuf<int> body_scene, row_scene;
  sfgs::DBuf<int> bscene_tab, bwg_scene, bgptr, bgslice;
  sfgs::DBuf<unsigned> bbar;
  sfgs::DBuf<double> bpart;
  sfgs::DBuf<unsigned char> bout;         // per-scene result records
  sfgs::DBuf<int> babort;
  sfgs::DBuf<int> u_clist, u_sscene, u_sbound, u_scene, u_active;   // per-colour batch schedule
  sfgs::DBuf<double> u_spart;

  // ---- device-resident time stepping (sfgs_dom.hip): body shapes and state
  // beyond the assembly's, materials, contact points and detection scratch
  int dom = 0;                            // shapes uploaded for the current bodies
  sfgs::DBuf<int> shape, surf, planes;    // planes: plane body indices, ascending
  sfgs::DBuf<double> rad, plane, sca, rot0, bx, J, mass, aux, force, point;
  int nplanes = 0;
  double rmax = 0.0;                      // largest bounding radius of a convex body
  int nmat = 0;
  sfgs::DBuf<int> mat_s;                  // surface pairs (smaller id first), 2 per material
  sfgs::DBuf<double> mat_v;               // friction, restitution, cohesion, spring, dashpot, hpow, 6 per material
  double defmat[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 1.0};
  double dom_time = 0.0;                  // domain time: contacts created at 0 take their material's cohesion
  int dom_dynamic = 1;                    // rigid scheme of the device loop: 1 dynamic, 0 quasi-static
  sfgs::DBuf<double> geo, geo2, dpart;
  sfgs::DBuf<int> okf, okpos, ok2, okpos2, ok3, okpos3, cand_m, cand_s;
  sfgs::DBuf<int> t_master, t_slave;
  sfgs::DBuf<double> t_R, t_U;
  sfgs::DBuf<unsigned long long> hkeys;   // existing body pairs (open addressing)
  sfgs::DBuf<int> cell_cnt, cell_start, cell_body, cell_cur, ncand, candoff, cand_a, cand_b;
};
