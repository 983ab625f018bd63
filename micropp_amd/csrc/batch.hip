// batch.hip -- many RVEs of uneven size, geometry and material state per
// launch (gfx950).
//
// A batch holds B RVEs, one per macro Gauss point.  Every finite-element
// operation of the Newton / CG path is ONE launch over the whole batch: the
// launch's blocks are the concatenation of each RVE's blocks, and a
// block -> RVE table (one int per block) tells a block which RVE it works on
// and its index inside that RVE.  The per-RVE arithmetic is the single-RVE
// kernels' (kernels.hpp bodies, same block shapes, same fixed reduction
// order), so a batched RVE's results are bit-identical to solving it alone.
// RVEs switched off in the per-RVE activity mask return at once; converged
// CGs stop through their own device-side flag.
//
// Launch shapes (blocks per RVE):
//   row   nrow_pad / 256       rows / nodes / CG vectors / SpMV
//   elem  ceil(nelem / 256)    one thread per element
//   ke    ceil(items / 4)      one wave per element stiffness (phase mode: 3)
//   rve   1                    per-RVE finishing reductions
#include "kernels.hpp"
#include "batch.hpp"

namespace mpp {

#define MPP_BATCH_PROLOGUE(MAP)                 \
  const int rid = MAP[blockIdx.x];              \
  if (!act[rid]) return;                        \
  const BRve &R = rv[rid];

// ---------------- boundary values and vector moves ----------------
__global__ void __launch_bounds__(256) kb_set_bc(const BRve *rv, const int *map, const int *act, const int *blk0,
                                                 const double *eps) {
  MPP_BATCH_PROLOGUE(map)
  const int lb = blockIdx.x - blk0[rid];
  set_bc_node(R.g, lb * 256 + threadIdx.x, eps + (size_t)rid * 6, R.uw);
}

// which: 0 uw <- u_n, 1 u_k <- uw, 2 uw <- u_k, 3 uw <- 0
__global__ void __launch_bounds__(256) kb_move(const BRve *rv, const int *map, const int *act, const int *blk0,
                                               int which) {
  MPP_BATCH_PROLOGUE(map)
  const int row = (blockIdx.x - blk0[rid]) * 256 + threadIdx.x;
  if (row >= R.g.nrow) return;
  switch (which) {
    case 0: R.uw[row] = R.u_n[row]; break;
    case 1: R.u_k[row] = R.uw[row]; break;
    case 2: R.uw[row] = R.u_k[row]; break;
    default: R.uw[row] = 0.0; break;
  }
}

// history moves: what 0: vars_n <- 0 and vars_k <- vars_aux (first
// nonlinearity); what 1: vars_aux <- 0
__global__ void __launch_bounds__(256) kb_vars_move(const BRve *rv, const int *map, const int *act, const int *blk0,
                                                    int what) {
  MPP_BATCH_PROLOGUE(map)
  const size_t n = (size_t)R.g.nelem * 8 * NVAR;
  const size_t t0 = (size_t)(blockIdx.x - blk0[rid]) * 256 + threadIdx.x;
  const size_t stride = (size_t)R.nb_elem * 256;
  for (size_t t = t0; t < n; t += stride) {
    if (what == 0) {
      R.vars_n[t] = 0.0;
      R.vars_k[t] = R.vars_aux[t];
    } else {
      R.vars_aux[t] = 0.0;
    }
  }
}

// ---------------- assembly ----------------
__global__ void __launch_bounds__(256) kb_elem_ke(const BRve *rv, const int *map, const int *act, const int *blk0) {
  MPP_BATCH_PROLOGUE(map)
  elem_ke_body(R.g, blockIdx.x - blk0[rid], R.bmat, R.mats, R.et, R.uw, R.use_vars ? R.vars_n : nullptr,
               R.phase_mode, R.ke);
}

__global__ void __launch_bounds__(256) kb_gather_vals(const BRve *rv, const int *map, const int *act,
                                                      const int *blk0) {
  MPP_BATCH_PROLOGUE(map)
  gather_vals_row(R.g, (blockIdx.x - blk0[rid]) * 256 + threadIdx.x, R.et, R.ke, R.phase_mode, R.vals);
}

__global__ void __launch_bounds__(256) kb_elem_be(const BRve *rv, const int *map, const int *act, const int *blk0) {
  MPP_BATCH_PROLOGUE(map)
  elem_be_body(R.g, blockIdx.x - blk0[rid], R.bmat, R.mats, R.et, R.uw, R.use_vars ? R.vars_n : nullptr, R.be);
}

__global__ void __launch_bounds__(256) kb_gather_rhs(const BRve *rv, const int *map, const int *act,
                                                     const int *blk0) {
  MPP_BATCH_PROLOGUE(map)
  gather_rhs_body(R.g, blockIdx.x - blk0[rid], R.be, R.b, R.part);
}

// one block per RVE: ||b||
__global__ void __launch_bounds__(256) kb_norm_finish(const BRve *rv, const int *act, double *out) {
  const int rid = blockIdx.x;
  if (!act[rid]) return;
  const BRve &R = rv[rid];
  double t[1];
  seg_sum<1>(R.part, R.nb_row, t);
  if (threadIdx.x == 0) out[rid] = sqrt(t[0]);
}

// u = uw (which 0) or u_k (which 1); vars as the RVE's use_vars says
__global__ void __launch_bounds__(256) kb_ave_stress(const BRve *rv, const int *map, const int *act, const int *blk0,
                                                     int which) {
  MPP_BATCH_PROLOGUE(map)
  ave_stress_body(R.g, blockIdx.x - blk0[rid], R.bmat, R.mats, R.et, which ? R.u_k : R.uw,
                  R.use_vars ? R.vars_n : nullptr, R.part);
}

__global__ void __launch_bounds__(256) kb_ave_finish(const BRve *rv, const int *act, double *out) {
  const int rid = blockIdx.x;
  if (!act[rid]) return;
  const BRve &R = rv[rid];
  double t[6];
  seg_sum<6>(R.part, R.nb_elem, t);
  if (threadIdx.x == 0)
    for (int v = 0; v < 6; ++v) out[(size_t)rid * 6 + v] = t[v] / R.g.vol_tot;
}

// calc_vars_new on u_k: vars_old = vars_n if the RVE has history, else none;
// vars_new = vars_k if it has history, else vars_aux.  flags[rid] |= nonlinear.
__global__ void __launch_bounds__(256) kb_vars_new(const BRve *rv, const int *map, const int *act, const int *blk0,
                                                   int *flags) {
  MPP_BATCH_PROLOGUE(map)
  vars_new_body(R.g, blockIdx.x - blk0[rid], R.bmat, R.mats, R.et, R.u_k, R.use_vars ? R.vars_n : nullptr,
                R.use_vars ? R.vars_k : R.vars_aux, flags + rid);
}

__global__ void kb_zero_ints(int n, int *p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// ---------------- CG ----------------
__device__ __forceinline__ const double *cg_matrix(const BRve &R, const int *a0on, int rid) {
  return a0on[rid] && R.a0 ? R.a0 : R.vals;
}

__global__ void __launch_bounds__(256) kb_cg_init(const BRve *rv, const int *map, const int *act, const int *blk0,
                                                  const int *a0on) {
  MPP_BATCH_PROLOGUE(map)
  cg_init_body(R.g, blockIdx.x - blk0[rid], cg_matrix(R, a0on, rid), R.b, R.x, R.r, R.z, R.p, R.kinv, R.part);
}

// one block per RVE; an inactive RVE's CG starts (and stays) done
__global__ void __launch_bounds__(256) kb_cg_init_state(const BRve *rv, const int *act, double min_err,
                                                        double rel_err, int max_its) {
  const int rid = blockIdx.x;
  const BRve &R = rv[rid];
  if (!act[rid]) {
    if (threadIdx.x == 0) {
      CgState c{};
      c.done = 1;
      R.cg[0] = c;
    }
    return;
  }
  cg_init_state_body(R.part, R.nb_row, min_err, rel_err, max_its, R.cg);
}

// The SpMV of the batch in one of its tuned shapes: CH slots of a row in
// flight per load round; XCD: launch block b takes logical block
// (b & 7) * per + (b >> 3), so each XCD streams a contiguous run of blocks
// (blocks are dispatched round-robin over the 8 XCDs).  The arithmetic is
// the same in every shape.
template <int CH, bool XCD>
__global__ void __launch_bounds__(256) kb_cg_step1(const BRve *rv, const int *map, const int *blk0, const int *a0on,
                                                   int nb, int per, int parity) {
  const int b = XCD ? (int)((blockIdx.x & 7) * per + (blockIdx.x >> 3)) : (int)blockIdx.x;
  if (b >= nb) return;
  const int rid = map[b];
  const BRve &R = rv[rid];
  // both matrices are assembled here: the stencil columns follow from the grid
  cg_step1_body<CH, true>(R.g, b - blk0[rid], nullptr, cg_matrix(R, a0on, rid), R.p, R.Ap, R.part,
                          R.cg + (parity & 1));
}

__global__ void __launch_bounds__(256) kb_cg_step2(const BRve *rv, const int *map, const int *blk0, int parity) {
  const int rid = map[blockIdx.x];
  const BRve &R = rv[rid];
  cg_step2_body(blockIdx.x - blk0[rid], R.nb_row, R.x, R.r, R.z, R.p, R.Ap, R.kinv, R.part, R.part + R.nb_row,
                R.cg + (parity & 1));
}

__global__ void __launch_bounds__(256) kb_cg_step3(const BRve *rv, const int *map, const int *blk0, int parity) {
  const int rid = map[blockIdx.x];
  const BRve &R = rv[rid];
  cg_step3_body(blockIdx.x - blk0[rid], R.nb_row, R.z, R.p, R.part, R.part + R.nb_row, R.cg + (parity & 1),
                R.cg + ((parity + 1) & 1));
}

// uw += x
__global__ void __launch_bounds__(256) kb_axpy(const BRve *rv, const int *map, const int *act, const int *blk0) {
#pragma clang fp contract(off)
  MPP_BATCH_PROLOGUE(map)
  const int row = (blockIdx.x - blk0[rid]) * 256 + threadIdx.x;
  if (row < R.g.nrow) R.uw[row] += 1.0 * R.x[row];
}

// ---------------- device-side Newton control ----------------
__device__ __forceinline__ void nt_test(BNewton &s, const BNewtonTest &t) {
  // the reference tests at the top of an iteration: at the cap it stops
  // unconverged, even when the last residual passes (src/solve.cpp:43-78)
  if (s.its >= t.max_its) {
    s.run = 0;
  } else if (s.norm < t.max_tol || s.norm < s.norm0 * t.rel_tol) {
    s.conv = 1;
    s.run = 0;
  }
}

__global__ void kb_nt_start(int n, const int *on, const double *norms, BNewtonTest t, BNewton *nst) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  BNewton s = nst[r];
  s.its = s.sits = s.conv = s.applied = 0;
  s.run = on[r] != 0;
  if (s.run) {
    s.norm0 = s.norm = norms[r];
    nt_test(s, t);
  }
  nst[r] = s;
}

__global__ void kb_nt_prepare(int n, const BRve *rv, BNewton *nst, int has_a0, int its_with_A0, int *act, int *ke,
                              int *gather, int *a0on, int *pend) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  BNewton s = nst[r];
  const int run = s.run;
  const int a0sel = run && has_a0 && s.its < its_with_A0;
  const int phase = rv[r].phase_mode;
  act[r] = run;
  if (has_a0) a0on[r] = a0sel;
  ke[r] = run && !a0sel && !phase;
  gather[r] = run && !a0sel && (!phase || !s.vals_ready);
  if (run && !a0sel && phase) s.vals_ready = 1;
  s.applied = 0;
  pend[r] = 0;
  nst[r] = s;
}

// rows: uw += x of the RVEs whose CG stopped and whose correction is not in uw
__global__ void __launch_bounds__(256) kb_nt_axpy(const BRve *rv, const int *map, const int *blk0,
                                                  const BNewton *nst) {
#pragma clang fp contract(off)
  const int rid = map[blockIdx.x];
  const BNewton &s = nst[rid];
  const BRve &R = rv[rid];
  if (!s.run || s.applied || !R.cg[0].done) return;
  const int row = (blockIdx.x - blk0[rid]) * 256 + threadIdx.x;
  if (row < R.g.nrow) R.uw[row] += 1.0 * R.x[row];
}

__global__ void kb_nt_mark(int n, const BRve *rv, BNewton *nst, int *pend) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  BNewton s = nst[r];
  const CgState c = rv[r].cg[0];
  const int now = s.run && !s.applied && c.done;
  if (now) {
    s.applied = 1;
    s.sits += c.its;
    nst[r] = s;
  }
  pend[r] = now;
}

__global__ void kb_nt_after(int n, const int *pend, const double *norms, BNewtonTest t, BNewton *nst) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n || !pend[r]) return;
  BNewton s = nst[r];
  s.its += 1;
  s.norm = norms[r];
  nt_test(s, t);
  nst[r] = s;
}

// ---------------- mixed precision ----------------
__global__ void __launch_bounds__(256) kb_m_prep(const BRve *rv, const int *map, const int *act, const int *blk0,
                                                 const int *a0on) {
  MPP_BATCH_PROLOGUE(map)
  const int row = (blockIdx.x - blk0[rid]) * 256 + threadIdx.x;  // < nrow_pad
  const double *A = cg_matrix(R, a0on, rid);
  const size_t base = (size_t)(row >> 6) * kSlots * 64 + (row & 63);
  for (int s = 0; s < kSlots; ++s) R.vals32[base + (size_t)s * 64] = (float)A[base + (size_t)s * 64];
  double k = 0.0;
  if (row < R.g.nrow) k = 1 / A[base + (size_t)(13 * 3 + row % 3) * 64];
  R.kinv[row] = k;
  R.kinv32[row] = (float)k;
  R.x[row] = 0.0;
}

// partials per row block: (D^-1 r)^2, r32.z32, z32.z32
__global__ void __launch_bounds__(256) kb_m_resid(const BRve *rv, const int *map, const int *act, const int *blk0,
                                                  const int *a0on) {
#pragma clang fp contract(off)
  __shared__ double lds[4 * 3];
  MPP_BATCH_PROLOGUE(map)
  const int lb = blockIdx.x - blk0[rid];
  const int row = lb * 256 + threadIdx.x;
  const double rr = R.b[row] - row_dot_grid(R.g, cg_matrix(R, a0on, rid), (const double *)R.x, row);
  const double zr = R.kinv[row] * rr;
  const float r32 = (float)rr;
  const float z32 = R.kinv32[row] * r32;
  R.r[row] = rr;
  R.r32[row] = r32;
  R.z32[row] = z32;
  R.p32[row] = z32;
  R.x32[row] = 0.0f;
  double v[3] = {zr * zr, (double)r32 * z32, (double)z32 * z32};
  block_sum<3, 256>(v, lds);
  if (threadIdx.x == 0)
    for (int q = 0; q < 3; ++q) R.part[(size_t)lb * 3 + q] = v[q];
}

// one block per RVE: the outer test and the inner CG's start state (an
// inactive or converged RVE's inner CG starts done)
__global__ void __launch_bounds__(256) kb_m_outer(const BRve *rv, const int *act, int first, double rel,
                                                  double abs_tol, double tau_min, int max_inner, const int *tight,
                                                  double tight_factor, BOuter *out) {
  const int rid = blockIdx.x;
  if (tight && tight[rid]) rel *= tight_factor;
  const BRve &R = rv[rid];
  if (!act[rid]) {
    if (threadIdx.x == 0) {
      CgState c{};
      c.done = 1;
      R.cg[0] = c;
    }
    return;
  }
  double t[3];
  seg_sum<3>(R.part, R.nb_row, t);
  if (threadIdx.x == 0) {
    BOuter &o = out[rid];
    const double pn = sqrt(t[0]);
    const double target = first ? fmax(abs_tol, rel * pn) : o.target;
    const int conv = pn < target || pn == 0.0;
    if (first) {
      o.pn0 = pn;
      o.target = target;
    }
    o.pn = pn;
    o.conv = conv;
    CgState c{};
    c.rz = t[1];
    c.pn0 = sqrt(t[2]);
    c.pn = c.pn0;
    c.min_err = 0.0;
    // aim a little below what the outer test needs, but not below what FP32 resolves
    c.rel_err = pn > 0.0 ? fmax(tau_min, kInnerAim * target / pn) : 1.0;
    c.max_its = max_inner;
    c.its = 0;
    c.done = conv || cg_done(c);
    R.cg[0] = c;
  }
}

// shapes as kb_cg_step1
template <int CH, bool XCD>
__global__ void __launch_bounds__(256) kb_m_step1(const BRve *rv, const int *map, const int *blk0, int nb, int per,
                                                  int parity) {
  const int b = XCD ? (int)((blockIdx.x & 7) * per + (blockIdx.x >> 3)) : (int)blockIdx.x;
  if (b >= nb) return;
  const int rid = map[b];
  const BRve &R = rv[rid];
  cg_step1_body<CH, true>(R.g, b - blk0[rid], nullptr, (const float *)R.vals32, (const float *)R.p32, R.Ap32,
                          R.part, R.cg + (parity & 1));
}

__global__ void __launch_bounds__(256) kb_m_step2(const BRve *rv, const int *map, const int *blk0, int parity) {
  const int rid = map[blockIdx.x];
  const BRve &R = rv[rid];
  cg_step2_body(blockIdx.x - blk0[rid], R.nb_row, R.x32, R.r32, R.z32, (const float *)R.p32, (const float *)R.Ap32,
                (const float *)R.kinv32, R.part, R.part + R.nb_row, R.cg + (parity & 1));
}

__global__ void __launch_bounds__(256) kb_m_step3(const BRve *rv, const int *map, const int *blk0, int parity) {
  const int rid = map[blockIdx.x];
  const BRve &R = rv[rid];
  cg_step3_body(blockIdx.x - blk0[rid], R.nb_row, (const float *)R.z32, R.p32, R.part, R.part + R.nb_row,
                R.cg + (parity & 1), R.cg + ((parity + 1) & 1));
}

__global__ void __launch_bounds__(256) kb_m_accum(const BRve *rv, const int *map, const int *act, const int *blk0) {
  MPP_BATCH_PROLOGUE(map)
  const int row = (blockIdx.x - blk0[rid]) * 256 + threadIdx.x;
  R.x[row] += (double)R.x32[row];
}

// ---------------------------------------------------------------------------
// Two-level preconditioner of the inner FP32 CG (mixed precision): z = D^-1 r
// + P_c A_c^-1 P_c^T r per RVE, with the trilinear hats of the RVE's coarse
// grid (R.cgrid) and the inverse of the Galerkin coarse matrix of its initial
// (elastic) tangent.  It runs at the inner CG's start, after kb_m_resid (act
// != nullptr: the RVEs in act), and in every iteration between kb_m_step2 and
// kb_m_step3 (act == nullptr: the RVEs whose CG state `parity` is running).
// FP32 residual rows: row = node * 3 + d.  The stopping test keeps the
// Jacobi ||D^-1 r|| of kb_m_step2; only (r, z) and z change.  Sums run in
// fixed orders.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool kb_c_skip(const BRve &R, const int *act, int rid, int parity) {
  return R.cgrid.nc == 0 || (act ? !act[rid] : R.cg[parity & 1].done != 0);
}

// restriction along x: one thread per (direction d, interior x-line (j, k),
// interior coarse x node cx), over the 2 hx - 1 nodes of the hat
__global__ void __launch_bounds__(256) kb_c_rx(const BRve *rv, const int *map, const int *blk0, const int *act,
                                               int parity) {
  const int rid = map[blockIdx.x];
  const BRve &R = rv[rid];
  if (kb_c_skip(R, act, rid, parity)) return;
  const Grid &g = R.g;
  const CGrid &cg = R.cgrid;
  const int ly = g.ny - 2, lz = g.nz - 2;
  const int t = (blockIdx.x - blk0[rid]) * 256 + threadIdx.x;
  if (t >= 3 * ly * lz * cg.mx) return;
  const int cx = t % cg.mx + 1;
  int rest = t / cg.mx;
  const int j = rest % ly + 1;
  rest /= ly;
  const int k = rest % lz + 1, d = rest / lz;
  const float *line = R.r32 + ((size_t)k * g.ny + j) * g.nx * 3 + d;
  const int i0 = (cx - 1) * cg.hx + 1, n = 2 * cg.hx - 1;
  float acc = 0.0f;
#pragma unroll 8
  for (int s = 0; s < n; ++s) acc += hat(i0 + s, cx, cg.hx) * line[(size_t)(i0 + s) * 3];
  R.ct1[((size_t)(d * g.nz + k) * g.ny + j) * cg.mx + (cx - 1)] = acc;
}

// restriction along y and z: one wave per coarse DOF q * 3 + d, its lanes over
// the x-restricted lines under q's hats, then a wave sum
__global__ void __launch_bounds__(256) kb_c_ryz(const BRve *rv, const int *map, const int *blk0, const int *act,
                                                int parity) {
  const int rid = map[blockIdx.x];
  const BRve &R = rv[rid];
  if (kb_c_skip(R, act, rid, parity)) return;
  const Grid &g = R.g;
  const CGrid &cg = R.cgrid;
  const int wave = (blockIdx.x - blk0[rid]) * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (wave >= cg.nc) return;  // whole wave
  const int q = wave / 3, d = wave % 3;
  const int cx = q % cg.mx + 1, cy = (q / cg.mx) % cg.my + 1, cz = q / (cg.mx * cg.my) + 1;
  const int wy = 2 * cg.hy - 1, npair = wy * (2 * cg.hz - 1);
  const int j0 = (cy - 1) * cg.hy + 1, k0 = (cz - 1) * cg.hz + 1;
  float acc = 0.0f;
  for (int s = lane; s < npair; s += 64) {
    const int j = j0 + s % wy, k = k0 + s / wy;
    acc += hat(j, cy, cg.hy) * hat(k, cz, cg.hz) * R.ct1[((size_t)(d * g.nz + k) * g.ny + j) * cg.mx + (cx - 1)];
  }
  const float sum = wave_sum_dpp(acc);
  if (lane == 0) R.crc[wave] = sum;
}

// the coarse solve e = A_c^-1 r_c: block (64 coarse rows, column split s of
// R.csplit), each wave a quarter of the split's columns (the symmetric
// inverse read by column), the waves added in order; into cec directly with
// one split, else into the split's partial row of cecp
__global__ void __launch_bounds__(256) kb_c_gemv(const BRve *rv, const int *map, const int *blk0, const int *act,
                                                 int parity) {
  __shared__ float red[3][64];
  const int rid = map[blockIdx.x];
  const BRve &R = rv[rid];
  if (kb_c_skip(R, act, rid, parity)) return;
  const int nc = R.cgrid.nc, lane = threadIdx.x & 63, ns = R.csplit;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lb = blockIdx.x - blk0[rid], tile = lb / ns, split = lb % ns;
  const int chunk = (nc + ns - 1) / ns, c0 = split * chunk, c1 = min(nc, c0 + chunk);
  const int row = tile * 64 + lane, rowc = min(row, nc - 1);
  const float *ainv = R.cainv, *rc = R.crc;
  float acc = 0.0f;
#pragma unroll 4
  for (int col = c0 + w; col < c1; col += 4) acc += ainv[(size_t)col * nc + rowc] * rc[col];
  if (w > 0) red[w - 1][lane] = acc;
  __syncthreads();
  if (w == 0 && row < nc) {
    const float v = acc + red[0][lane] + red[1][lane] + red[2][lane];
    if (ns == 1)
      R.cec[row] = v;
    else
      R.cecp[(size_t)split * nc + row] = v;
  }
}

// the splits' sum in order (RVEs with csplit > 1)
__global__ void __launch_bounds__(256) kb_c_gsum(const BRve *rv, const int *map, const int *blk0, const int *act,
                                                 int parity) {
  const int rid = map[blockIdx.x];
  const BRve &R = rv[rid];
  if (kb_c_skip(R, act, rid, parity)) return;
  const int nc = R.cgrid.nc, row = (blockIdx.x - blk0[rid]) * 256 + threadIdx.x;
  if (row >= nc) return;
  float v = 0.0f;
  for (int s = 0; s < R.csplit; ++s) v += R.cecp[(size_t)s * nc + row];
  R.cec[row] = v;
}

// z += P_c e on the RVE's rows (the 8 coarse nodes around a node, zero
// weights outside) and the block's (r, z) partial in place of the Jacobi one:
// part[lb * 3 + 1] at the start (where p = z as well), else the (r, z) slot
// of kb_m_step2's partials
__global__ void __launch_bounds__(256) kb_c_zfix(const BRve *rv, const int *map, const int *blk0, const int *act,
                                                 int parity) {
  __shared__ double lds[4];
  const int rid = map[blockIdx.x];
  const BRve &R = rv[rid];
  if (kb_c_skip(R, act, rid, parity)) return;
  const Grid &g = R.g;
  const CGrid &cg = R.cgrid;
  const int lb = blockIdx.x - blk0[rid], row = lb * 256 + threadIdx.x;  // < nrow_pad
  float z = R.z32[row];
  const float r = R.r32[row];
  if (row < g.nrow) {
    const int node = row / 3, d = row % 3;
    const int i = node % g.nx, j = (node / g.nx) % g.ny, k = node / (g.nx * g.ny);
    float wx[2], wy[2], wz[2];
    int qx[2], qy[2], qz[2];
    hat_pair(i, cg.hx, cg.mx, wx, qx);
    hat_pair(j, cg.hy, cg.my, wy, qy);
    hat_pair(k, cg.hz, cg.mz, wz, qz);
    float e[8];
#pragma unroll
    for (int t = 0; t < 8; ++t)
      e[t] = R.cec[((qz[t >> 2] * cg.my + qy[(t >> 1) & 1]) * cg.mx + qx[t & 1]) * 3 + d];
    float acc = 0.0f;
#pragma unroll
    for (int t = 0; t < 8; ++t) acc += wx[t & 1] * (wy[(t >> 1) & 1] * wz[t >> 2]) * e[t];
    z += acc;
  }
  R.z32[row] = z;
  if (act) R.p32[row] = z;
  double v[1] = {(double)r * z};
  block_sum<1, 256>(v, lds);
  if (threadIdx.x == 0) {
    if (act)
      R.part[(size_t)lb * 3 + 1] = v[0];
    else
      R.part[R.nb_row + (size_t)lb * 2 + 1] = v[0];
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static hipError_t launch_bc(const BLaunch &L, const int *act, int parity, hipStream_t st) {
  hipLaunchKernelGGL(kb_c_rx, dim3(L.crx.nb), dim3(256), 0, st, L.rv, L.crx.map, L.crx.blk0, act, parity);
  hipLaunchKernelGGL(kb_c_ryz, dim3(L.cryz.nb), dim3(256), 0, st, L.rv, L.cryz.map, L.cryz.blk0, act, parity);
  hipLaunchKernelGGL(kb_c_gemv, dim3(L.cgemv.nb), dim3(256), 0, st, L.rv, L.cgemv.map, L.cgemv.blk0, act, parity);
  if (L.cgsum.nb)
    hipLaunchKernelGGL(kb_c_gsum, dim3(L.cgsum.nb), dim3(256), 0, st, L.rv, L.cgsum.map, L.cgsum.blk0, act, parity);
  hipLaunchKernelGGL(kb_c_zfix, dim3(L.row.nb), dim3(256), 0, st, L.rv, L.row.map, L.row.blk0, act, parity);
  return hipGetLastError();
}

hipError_t launch_b_set_bc(const BLaunch &L, const double *eps, hipStream_t st) {
  hipLaunchKernelGGL(kb_set_bc, dim3(L.row.nb), dim3(256), 0, st, L.rv, L.row.map, L.act, L.row.blk0, eps);
  return hipGetLastError();
}
hipError_t launch_b_move(const BLaunch &L, int which, hipStream_t st) {
  hipLaunchKernelGGL(kb_move, dim3(L.row.nb), dim3(256), 0, st, L.rv, L.row.map, L.act, L.row.blk0, which);
  return hipGetLastError();
}
hipError_t launch_b_vars_move(const BLaunch &L, int what, hipStream_t st) {
  hipLaunchKernelGGL(kb_vars_move, dim3(L.elem.nb), dim3(256), 0, st, L.rv, L.elem.map, L.act, L.elem.blk0, what);
  return hipGetLastError();
}
hipError_t launch_b_mat(const BLaunch &L, const int *act_ke, const int *act_gather, hipStream_t st) {
  hipLaunchKernelGGL(kb_elem_ke, dim3(L.ke.nb), dim3(64 * kKeWaves), 0, st, L.rv, L.ke.map, act_ke, L.ke.blk0);
  hipLaunchKernelGGL(kb_gather_vals, dim3(L.row.nb), dim3(256), 0, st, L.rv, L.row.map, act_gather, L.row.blk0);
  return hipGetLastError();
}
hipError_t launch_b_rhs(const BLaunch &L, double *norms, hipStream_t st) {
  hipLaunchKernelGGL(kb_elem_be, dim3(L.elem.nb), dim3(256), 0, st, L.rv, L.elem.map, L.act, L.elem.blk0);
  hipLaunchKernelGGL(kb_gather_rhs, dim3(L.row.nb), dim3(256), 0, st, L.rv, L.row.map, L.act, L.row.blk0);
  hipLaunchKernelGGL(kb_norm_finish, dim3(L.nrve), dim3(256), 0, st, L.rv, L.act, norms);
  return hipGetLastError();
}
hipError_t launch_b_ave(const BLaunch &L, int which, double *sig, hipStream_t st) {
  hipLaunchKernelGGL(kb_ave_stress, dim3(L.elem.nb), dim3(256), 0, st, L.rv, L.elem.map, L.act, L.elem.blk0, which);
  hipLaunchKernelGGL(kb_ave_finish, dim3(L.nrve), dim3(256), 0, st, L.rv, L.act, sig);
  return hipGetLastError();
}
hipError_t launch_b_vars_new(const BLaunch &L, int *flags, hipStream_t st) {
  hipLaunchKernelGGL(kb_zero_ints, dim3((L.nrve + 255) / 256), dim3(256), 0, st, L.nrve, flags);
  hipLaunchKernelGGL(kb_vars_new, dim3(L.elem.nb), dim3(256), 0, st, L.rv, L.elem.map, L.act, L.elem.blk0, flags);
  return hipGetLastError();
}
hipError_t launch_b_cg_init(const BLaunch &L, double min_err, double rel_err, int max_its, hipStream_t st) {
  hipLaunchKernelGGL(kb_cg_init, dim3(L.row.nb), dim3(256), 0, st, L.rv, L.row.map, L.act, L.row.blk0, L.a0on);
  hipLaunchKernelGGL(kb_cg_init_state, dim3(L.nrve), dim3(256), 0, st, L.rv, L.act, min_err, rel_err, max_its);
  return hipGetLastError();
}
hipError_t launch_b_spmv(const BLaunch &L, int shape, int parity, hipStream_t st) {
  const int nb = L.row.nb, per = (nb + 7) / 8;
  const dim3 plain(nb), xcd(8 * per);
  switch (shape) {
    case 0: hipLaunchKernelGGL((kb_cg_step1<9, false>), plain, dim3(256), 0, st, L.rv, L.row.map, L.row.blk0, L.a0on, nb, per, parity); break;
    case 1: hipLaunchKernelGGL((kb_cg_step1<27, false>), plain, dim3(256), 0, st, L.rv, L.row.map, L.row.blk0, L.a0on, nb, per, parity); break;
    case 2: hipLaunchKernelGGL((kb_cg_step1<81, false>), plain, dim3(256), 0, st, L.rv, L.row.map, L.row.blk0, L.a0on, nb, per, parity); break;
    case 3: hipLaunchKernelGGL((kb_cg_step1<9, true>), xcd, dim3(256), 0, st, L.rv, L.row.map, L.row.blk0, L.a0on, nb, per, parity); break;
    case 4: hipLaunchKernelGGL((kb_cg_step1<27, true>), xcd, dim3(256), 0, st, L.rv, L.row.map, L.row.blk0, L.a0on, nb, per, parity); break;
    case 5: hipLaunchKernelGGL((kb_cg_step1<81, true>), xcd, dim3(256), 0, st, L.rv, L.row.map, L.row.blk0, L.a0on, nb, per, parity); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_b_cg_iter(const BLaunch &L, int parity, hipStream_t st) {
  hipError_t e = launch_b_spmv(L, L.spmv_shape, parity, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kb_cg_step2, dim3(L.row.nb), dim3(256), 0, st, L.rv, L.row.map, L.row.blk0, parity);
  hipLaunchKernelGGL(kb_cg_step3, dim3(L.row.nb), dim3(256), 0, st, L.rv, L.row.map, L.row.blk0, parity);
  return hipGetLastError();
}
hipError_t launch_b_axpy(const BLaunch &L, hipStream_t st) {
  hipLaunchKernelGGL(kb_axpy, dim3(L.row.nb), dim3(256), 0, st, L.rv, L.row.map, L.act, L.row.blk0);
  return hipGetLastError();
}

static unsigned per_rve_blocks(int n) { return (unsigned)((n + 255) / 256); }

hipError_t launch_b_nt_start(const BLaunch &L, const int *on, const double *norms, BNewtonTest t, BNewton *nst,
                             hipStream_t st) {
  hipLaunchKernelGGL(kb_nt_start, dim3(per_rve_blocks(L.nrve)), dim3(256), 0, st, L.nrve, on, norms, t, nst);
  return hipGetLastError();
}
hipError_t launch_b_nt_prepare(const BLaunch &L, BNewton *nst, int has_a0, int its_with_A0, int *act, int *ke,
                               int *gather, int *a0on, int *pend, hipStream_t st) {
  hipLaunchKernelGGL(kb_nt_prepare, dim3(per_rve_blocks(L.nrve)), dim3(256), 0, st, L.nrve, L.rv, nst, has_a0,
                     its_with_A0, act, ke, gather, a0on, pend);
  return hipGetLastError();
}
hipError_t launch_b_nt_finish(const BLaunch &L, BNewtonTest t, BNewton *nst, int *pend, double *norms,
                              hipStream_t st) {
  hipLaunchKernelGGL(kb_nt_axpy, dim3(L.row.nb), dim3(256), 0, st, L.rv, L.row.map, L.row.blk0, nst);
  hipLaunchKernelGGL(kb_nt_mark, dim3(per_rve_blocks(L.nrve)), dim3(256), 0, st, L.nrve, L.rv, nst, pend);
  BLaunch P = L;
  P.act = pend;
  hipError_t e = launch_b_rhs(P, norms, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kb_nt_after, dim3(per_rve_blocks(L.nrve)), dim3(256), 0, st, L.nrve, pend, norms, t, nst);
  return hipGetLastError();
}

hipError_t launch_bm_prep(const BLaunch &L, hipStream_t st) {
  hipLaunchKernelGGL(kb_m_prep, dim3(L.row.nb), dim3(256), 0, st, L.rv, L.row.map, L.act, L.row.blk0, L.a0on);
  return hipGetLastError();
}
hipError_t launch_bm_resid(const BLaunch &L, int first, double rel, double abs_tol, double tau_min, int max_inner,
                           const int *tight, double tight_factor, BOuter *out, hipStream_t st) {
  hipLaunchKernelGGL(kb_m_resid, dim3(L.row.nb), dim3(256), 0, st, L.rv, L.row.map, L.act, L.row.blk0, L.a0on);
  if (L.coarse) {
    const hipError_t e = launch_bc(L, L.act, 0, st);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(kb_m_outer, dim3(L.nrve), dim3(256), 0, st, L.rv, L.act, first, rel, abs_tol, tau_min,
                     max_inner, tight, tight_factor, out);
  return hipGetLastError();
}
hipError_t launch_bm_spmv(const BLaunch &L, int shape, int parity, hipStream_t st) {
  const int nb = L.row.nb, per = (nb + 7) / 8;
  const dim3 plain(nb), xcd(8 * per);
  switch (shape) {
    case 0: hipLaunchKernelGGL((kb_m_step1<9, false>), plain, dim3(256), 0, st, L.rv, L.row.map, L.row.blk0, nb, per, parity); break;
    case 1: hipLaunchKernelGGL((kb_m_step1<27, false>), plain, dim3(256), 0, st, L.rv, L.row.map, L.row.blk0, nb, per, parity); break;
    case 2: hipLaunchKernelGGL((kb_m_step1<81, false>), plain, dim3(256), 0, st, L.rv, L.row.map, L.row.blk0, nb, per, parity); break;
    case 3: hipLaunchKernelGGL((kb_m_step1<9, true>), xcd, dim3(256), 0, st, L.rv, L.row.map, L.row.blk0, nb, per, parity); break;
    case 4: hipLaunchKernelGGL((kb_m_step1<27, true>), xcd, dim3(256), 0, st, L.rv, L.row.map, L.row.blk0, nb, per, parity); break;
    case 5: hipLaunchKernelGGL((kb_m_step1<81, true>), xcd, dim3(256), 0, st, L.rv, L.row.map, L.row.blk0, nb, per, parity); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
hipError_t launch_bm_iter(const BLaunch &L, int parity, hipStream_t st) {
  const hipError_t e = launch_bm_spmv(L, L.spmv_shape32, parity, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kb_m_step2, dim3(L.row.nb), dim3(256), 0, st, L.rv, L.row.map, L.row.blk0, parity);
  if (L.coarse) {
    const hipError_t e2 = launch_bc(L, nullptr, parity, st);
    if (e2 != hipSuccess) return e2;
  }
  hipLaunchKernelGGL(kb_m_step3, dim3(L.row.nb), dim3(256), 0, st, L.rv, L.row.map, L.row.blk0, parity);
  return hipGetLastError();
}
hipError_t launch_bm_accum(const BLaunch &L, hipStream_t st) {
  hipLaunchKernelGGL(kb_m_accum, dim3(L.row.nb), dim3(256), 0, st, L.rv, L.row.map, L.act, L.row.blk0);
  return hipGetLastError();
}

}  // namespace mpp
