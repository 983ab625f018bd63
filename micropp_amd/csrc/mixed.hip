// mixed.hip -- mixed-precision multi-right-hand-side solver for one elastic
// RVE (gfx950).
//
// K affine-boundary problems on the same RVE (the six unit strains of ctan
// and the macro strains of sigma) are solved together.  The outer loop is an
// FP64 defect correction (a Newton iteration of the linear problem): the
// residual b - A u is computed matrix-free in FP64 from the element
// stiffness of each phase, and u += d is applied in FP64.  The correction d
// comes from an FP32 Jacobi-preconditioned CG on all K systems at once: one
// pass over the FP32 matrix serves K vectors.  The volume averages are FP64.
//
// The inner CG keeps the reference's recurrence (alpha from (p, Ap) computed
// directly: the single-reduction Chronopoulos-Gear form, tried here, loses
// that denominator to cancellation in FP32 after ~150 iterations and
// diverges), with the stopping test on ||z||, the reference CG's metric.
//
// FP32 matrix layout (node-blocked SELL-64): the 27 neighbour blocks of 3x3
// coefficients of a node (nb as in the reference ELL) in 9 stencil rows
// (dz, dy) of 3 neighbours (dx = -1, 0, 1); a row's 27 floats plus 1 pad are 7
// float4, so slot (nb, fi, fj) is float (nb / 3) * 28 + (nb % 3) * 9 + fi * 3
// + fj and a node has 63 float4, quad-major inside a slice of 64 nodes: quad
// q of node n sits at float4 index ((n / 64) * 63 + q) * 64 + n % 64.  Column
// indices are implicit (structured 27-point stencil).  Rows and columns of
// Dirichlet nodes are dropped (the CG vectors are zero there), which leaves
// the symmetric interior block.
//
// FP32 vectors group 4 fields: float4 (d * K/4 + g) * NP + n holds fields
// 4g .. 4g+3 of DOF direction d at node n (NP = nodes rounded up to 256).  A
// wave of the SpMV owns 64 consecutive nodes and one field group, so every
// matrix load and every neighbour gather is one contiguous 1 KB line.  The
// FP64 displacements U are per field: component c = d * K + k of node n at
// c * NP + n.
//
// Reductions are fence-free: blocks write FP64 partials in a [row][block]
// table, and each row is re-summed in a fixed order, by a one-block kernel or
// identically by every block of the consumer kernel.  Iteration
// k of the inner CG reads state[k & 1] and writes state[(k+1) & 1].  The
// SpMV maps slices to XCDs so that each XCD's L2 holds a contiguous slab of
// nodes.
#include "internal.hpp"
#include "kernels.hpp"
#include "physics.hpp"
#include "reduce.hpp"

namespace mpp {

constexpr int kMaxRhs = 16;
constexpr int kRowQuads = 7, kRows = 9;         // stencil rows (dz, dy) of 27 floats + 1 pad
constexpr int kNodeQuads = kRowQuads * kRows;  // 63 float4 per node
__host__ __device__ constexpr int mslot(int nb, int fi, int fj) { return (nb / 3) * 28 + (nb % 3) * 9 + fi * 3 + fj; }
constexpr int kRedThreads = 256;  // threads of the one-block reducers
constexpr int kVecNodes = 2048;   // nodes per block of the vector kernels

// inner CG scalars per field: rz = (r, z), alpha / beta of the iteration, pn
// the current ||z||, pn0 the start value, tau the relative stop
struct MState {
  double rz[kMaxRhs], alpha[kMaxRhs], beta[kMaxRhs], pn[kMaxRhs], pn0[kMaxRhs], tau[kMaxRhs];
  int done[kMaxRhs];
  int its, max_its, all_done, pad;
};

// outer (defect-correction) test per field: pn = ||kinv r|| (the reference
// CG's metric), rn = ||r||_2 over the interior DOFs (assembly_rhs's norm:
// its Dirichlet rows are zero), and their values at the last `first` call
struct MOuter {
  double pn[kMaxRhs], pn0[kMaxRhs], target[kMaxRhs], rn[kMaxRhs], rn0[kMaxRhs];
  int conv[kMaxRhs];
  int all_conv, pad[3];
};

typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bool node_on_boundary(const Grid &g, int i, int j, int k) {
  return i == 0 || j == 0 || k == 0 || i == g.nx - 1 || j == g.ny - 1 || k == g.nz - 1;
}

// Blocks are dispatched round-robin over the 8 XCDs; launch 8 * per blocks and
// give XCD x the logical blocks [x * per, (x + 1) * per).
__device__ __forceinline__ int xcd_block(int per) {
  const int b = blockIdx.x;
  return (b & 7) * per + (b >> 3);
}

// Row sums of a partial table part[row * stride + j], j < n, rows < NR, by
// one block of kRedThreads: every thread accumulates its entries j = t,
// t + kRedThreads, ... of all rows (independent loads, all in flight), each
// wave adds its lanes with DPP moves, and thread t < NR adds the waves' sums
// of row t.  Fixed order throughout.  Thread t < NR returns the sum of row t
// (others return 0).
// The two halves of rows_sum, for callers that issue other loads in between
// (vector-memory loads complete in issue order, so loads issued after the
// partials' do not delay their sum): rows_load adds this thread's entries of
// each row in index order, rows_finish combines the threads.
template <int NR, int UNROLL = 1>
__device__ __forceinline__ void rows_load(const double *part, size_t stride, int n, double (&s)[NR]) {
#pragma unroll
  for (int q = 0; q < NR; ++q) s[q] = 0.0;
  // UNROLL > 1 keeps several rounds of loads in flight (long rows)
#pragma unroll UNROLL
  for (int j = threadIdx.x; j < n; j += kRedThreads)
#pragma unroll
    for (int q = 0; q < NR; ++q) s[q] += part[(size_t)q * stride + j];
}
template <int NR>
__device__ __forceinline__ double rows_finish(double (&s)[NR]) {
  static_assert(NR <= kRedThreads, "one row per thread at the end");
  constexpr int kW = kRedThreads / 64;
  __shared__ double lds[kW][NR];
  const int t = threadIdx.x, w = t >> 6;
  __syncthreads();  // lds may still be read by a previous call
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    const double v = wave_sum_dpp(s[q]);
    if ((t & 63) == 0) lds[w][q] = v;
  }
  __syncthreads();
  double r = 0.0;
  if (t < NR)
    for (int i = 0; i < kW; ++i) r += lds[i][t];
  return r;
}
template <int NR, int UNROLL = 1>
__device__ __forceinline__ double rows_sum(const double *part, size_t stride, int n) {
  double s[NR];
  rows_load<NR, UNROLL>(part, stride, n, s);
  return rows_finish<NR>(s);
}

// ---------------------------------------------------------------------------
// The element blocks (phase table or per element) around interior node n:
// blk[q] the 24x24 block of element q = (a*2+b)*2+c, la[q] the node's local
// index in it.  Entry (fi, fj) of the node block towards neighbour nb is the
// FP64 sum of the element contributions in the reference element order.
// ---------------------------------------------------------------------------
struct NodeBlocks {
  const double *blk[8];
  int la[8];
  __device__ NodeBlocks(const Grid &g, int i, int j, int k, const int *et, const double *ke, int phase_mode) {
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        for (int c = 0; c < 2; ++c) {
          const int q = (a * 2 + b) * 2 + c;
          const int e = ((k - 1 + c) * g.ney + (j - 1 + b)) * g.nex + (i - 1 + a);
          blk[q] = ke + (size_t)(phase_mode ? et[e] : e) * 576;
          la[q] = local_node(1 - a, 1 - b, 1 - c);
        }
  }
  __device__ double entry(int nb, int fi, int fj) const {
    const int dz = nb / 9 - 1, dy = (nb / 3) % 3 - 1, dx = nb % 3 - 1;
    double acc = 0.0;
    for (int a = 0; a < 2; ++a) {
      const int px = 1 - a + dx;
      if (px < 0 || px > 1) continue;
      for (int b = 0; b < 2; ++b) {
        const int py = 1 - b + dy;
        if (py < 0 || py > 1) continue;
        for (int c = 0; c < 2; ++c) {
          const int pz = 1 - c + dz;
          if (pz < 0 || pz > 1) continue;
          const int q = (a * 2 + b) * 2 + c;
          acc += blk[q][(la[q] * 3 + fi) * 24 + local_node(px, py, pz) * 3 + fj];
        }
      }
    }
    return acc;
  }
};

// Jacobi weights: kinv64[d * NP + n] = 1 / A_nn(d, d) (FP64, zero on
// Dirichlet DOFs and padding) and its FP32 copy kinv
__global__ void __launch_bounds__(256) k_m_kinv(Grid g, int NP, const int *et, const double *ke, int phase_mode,
                                                double *kinv64, float *kinv) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= NP) return;
  double w[3] = {0.0, 0.0, 0.0};
  if (n < g.nn) {
    const int i = n % g.nx, j = (n / g.nx) % g.ny, k = n / (g.nx * g.ny);
    if (!node_on_boundary(g, i, j, k)) {
      const NodeBlocks B(g, i, j, k, et, ke, phase_mode);
      for (int d = 0; d < 3; ++d) w[d] = 1.0 / B.entry(13, d, d);
    }
  }
  for (int d = 0; d < 3; ++d) {
    kinv64[(size_t)d * NP + n] = w[d];
    kinv[(size_t)d * NP + n] = (float)w[d];
  }
}

// FP32 matrix: slot (nb, fi, fj) of node n = A(n fi, m fj) for neighbour m,
// rounded once; rows and columns of Dirichlet DOFs are zero
__global__ void __launch_bounds__(256) k_m_asm32(Grid g, int NP, const int *et, const double *ke, int phase_mode,
                                                 float *A) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= NP) return;
  float *base = A + ((size_t)(n >> 6) * kNodeQuads * 64 + (n & 63)) * 4;
  auto put = [&](int s, float v) { base[(size_t)(s >> 2) * 256 + (s & 3)] = v; };
  for (int s = 0; s < kNodeQuads * 4; ++s) put(s, 0.0f);
  if (n >= g.nn) return;
  const int i = n % g.nx, j = (n / g.nx) % g.ny, k = n / (g.nx * g.ny);
  if (node_on_boundary(g, i, j, k)) return;
  const NodeBlocks B(g, i, j, k, et, ke, phase_mode);
  for (int nb = 0; nb < 27; ++nb) {
    const int dz = nb / 9 - 1, dy = (nb / 3) % 3 - 1, dx = nb % 3 - 1;
    if (node_on_boundary(g, i + dx, j + dy, k + dz)) continue;  // column of a Dirichlet DOF: dropped
    for (int fi = 0; fi < 3; ++fi)
      for (int fj = 0; fj < 3; ++fj) put(mslot(nb, fi, fj), (float)B.entry(nb, fi, fj));
  }
}

// ---------------------------------------------------------------------------
// boundary values of K fields (set_displ_bc, src/micro3D.cpp:26-78), FP64
// ---------------------------------------------------------------------------
template <int K>
__global__ void __launch_bounds__(256) k_m_setbc(Grid g, int NP, const double *eps, double *U) {
#pragma clang fp contract(off)
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.nn) return;
  const int i = n % g.nx, j = (n / g.nx) % g.ny, k = n / (g.nx * g.ny);
  if (!node_on_boundary(g, i, j, k)) return;
  const double L = 1.0;
  double x[3];
  if (i == g.nx - 1) {
    x[0] = L, x[1] = j * g.dy, x[2] = k * g.dz;
  } else if (i == 0) {
    x[0] = 0, x[1] = j * g.dy, x[2] = k * g.dz;
  } else if (j == g.ny - 1) {
    x[0] = i * g.dx, x[1] = L, x[2] = k * g.dz;
  } else if (j == 0) {
    x[0] = i * g.dx, x[1] = 0, x[2] = k * g.dz;
  } else if (k == g.nz - 1) {
    x[0] = i * g.dx, x[1] = j * g.dy, x[2] = L;
  } else {
    x[0] = i * g.dx, x[1] = j * g.dy, x[2] = 0;
  }
  for (int r = 0; r < K; ++r) {
    const double *e = eps + r * 6;
    const double et[3][3] = {{e[0], 0.5 * e[3], 0.5 * e[4]}, {0.5 * e[3], e[1], 0.5 * e[5]}, {0.5 * e[4], 0.5 * e[5], e[2]}};
    for (int d = 0; d < 3; ++d) {
      double s = 0.0;
      for (int c = 0; c < 3; ++c) s += et[d][c] * x[c];
      U[(size_t)(d * K + r) * NP + n] = s;
    }
  }
}

// ---------------------------------------------------------------------------
// One field of U from / into a displacement vector in the reference DOF
// numbering (u[node * 3 + d]); padding nodes of U are zero after a load.
// ---------------------------------------------------------------------------
template <int K>
__global__ void __launch_bounds__(256) k_m_load_u(Grid g, int NP, int f, const double *u, double *U) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= NP) return;
  for (int d = 0; d < 3; ++d) U[(size_t)(d * K + f) * NP + n] = (n < g.nn) ? u[(size_t)n * 3 + d] : 0.0;
}

template <int K>
__global__ void __launch_bounds__(256) k_m_store_u(Grid g, int NP, int f, const double *U, double *u) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.nn) return;
  for (int d = 0; d < 3; ++d) u[(size_t)n * 3 + d] = U[(size_t)(d * K + f) * NP + n];
}

// ---------------------------------------------------------------------------
// Adjoint sources.  For sigma_v = (1/V) sum_e sum_gp wg [C B u_e]_v the
// derivative with respect to the displacements is assembled from, per phase
// p, G[p][q][v] = wg sum_m C_p[v][m] sum_gp B_gp[m][q] (q < 24 element DOFs).
// The adjoint field v solves K_II lambda_v = (1/V) sum_e G[phase(e)][.][v] on
// the interior DOFs, with lambda_v = 0 on the boundary.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_m_adj_G(const double *bmat, const material_base *mats, double wg,
                                                 double *G) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * 24 * 6) return;
  const int p = t / 144, q = (t / 6) % 24, v = t % 6;
  double s = 0.0;
  for (int m = 0; m < 6; ++m) {
    double e[6] = {0, 0, 0, 0, 0, 0}, sg[6];
    e[m] = 1.0;
    lin_stress(mats[p], e, sg);  // sg[v] = C[v][m]
    double bs = 0.0;
    for (int gp = 0; gp < 8; ++gp) bs += bmat[gp * 144 + m * 24 + q];
    s += sg[v] * bs;
  }
  G[t] = wg * s;
}

// F[(d K + v) NP + n]: the adjoint source of field v < nadj at interior node n;
// zero on the boundary, on the padding and for the fields v >= nadj
template <int K>
__global__ void __launch_bounds__(256) k_m_adj_src(Grid g, int NP, int nadj, const int *et, const double *G,
                                                   double *F) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= NP) return;
  double f[3][K];
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int v = 0; v < K; ++v) f[d][v] = 0.0;
  if (n < g.nn) {
    const int i = n % g.nx, j = (n / g.nx) % g.ny, k = n / (g.nx * g.ny);
    if (!node_on_boundary(g, i, j, k)) {
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
          for (int c = 0; c < 2; ++c) {
            const int e = ((k - 1 + c) * g.ney + (j - 1 + b)) * g.nex + (i - 1 + a);
            const double *Gp = G + (size_t)et[e] * 144 + local_node(1 - a, 1 - b, 1 - c) * 3 * 6;
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
              for (int v = 0; v < 6 && v < K; ++v)
                if (v < nadj) f[d][v] += Gp[d * 6 + v];
          }
#pragma unroll
      for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int v = 0; v < K; ++v) f[d][v] /= g.vol_tot;
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int v = 0; v < K; ++v) F[(size_t)(d * K + v) * NP + n] = f[d][v];
}

// ---------------------------------------------------------------------------
// FP64 residual r = F - (K u) on interior DOFs (elastic: element blocks of the
// phase table, summed per stencil neighbour; F = 0 unless adjoint sources are
// given), rounded into the FP32 CG start: R = r, X = 0, P = kinv R.
// Partials: part[q * gridDim.x + block], q < K: ||kinv r||^2 (FP64) of field
// q; K <= q < 2K: (R, P) of field q - K; 2K <= q: ||r||^2 (FP64) of field
// q - 2K.  Grid (node blocks, K/4): block (b, grp) takes fields 4 grp ..
// 4 grp + 3 (each group rebuilds the node's stencil; four fields' sums keep
// the registers at two waves per SIMD, twelve took all 256 VGPRs).
// ---------------------------------------------------------------------------
template <int K>
__global__ void __launch_bounds__(256) k_m_resid(Grid g, int NP, int per, const int *et, const double *ke,
                                                 const double *U, const double *F, const float *kinv, float4 *R,
                                                 float4 *X, float4 *P, double *part) {
  constexpr int KG = K / 4;
  __shared__ double ke_s[3 * 576];
  __shared__ double lds[4 * 3 * 4];
  const int grp = blockIdx.y, f0 = 4 * grp;
  for (int t = threadIdx.x; t < 3 * 576; t += blockDim.x) ke_s[t] = ke[t];
  __syncthreads();
  const int lb = xcd_block(per);
  const int n = lb * 256 + threadIdx.x;
  double zz[4], rz[4], rr[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) zz[r] = rz[r] = rr[r] = 0.0;
  if (n < NP) {
    const int i = n % g.nx, j = (n / g.nx) % g.ny, k = n / (g.nx * g.ny);
    const bool interior = n < g.nn && !node_on_boundary(g, i, j, k);
    // (K u) at node n by stencil neighbour: the 3x3 block towards neighbour
    // m = n + (dx, dy, dz) is the sum of the element blocks Ke(n, m) over the
    // (1, 2, 4 or 8) elements holding both, and is applied to every field
    // once, so each neighbour's u is loaded once per field.
    double acc[3][4];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[d][q] = 0.0;
    if (interior) {
      const double *kbe[8];  // element (a, b, c) = (i - 1 + a, j - 1 + b, k - 1 + c): its phase block
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
          for (int a = 0; a < 2; ++a)
            kbe[(c * 2 + b) * 2 + a] = ke_s + et[((k - 1 + c) * g.ney + (j - 1 + b)) * g.nex + (i - 1 + a)] * 576;
#pragma unroll
      for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
          for (int dx = -1; dx <= 1; ++dx) {
            double cf[9];
#pragma unroll
            for (int t = 0; t < 9; ++t) cf[t] = 0.0;
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
              for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int a = 0; a < 2; ++a) {
                  const int mx = 1 - a + dx, my = 1 - b + dy, mz = 1 - c + dz;  // m in element (a, b, c)
                  if (mx < 0 || mx > 1 || my < 0 || my > 1 || mz < 0 || mz > 1) continue;
                  const double *kb = kbe[(c * 2 + b) * 2 + a] + local_node(1 - a, 1 - b, 1 - c) * 3 * 24 +
                                     local_node(mx, my, mz) * 3;
#pragma unroll
                  for (int fi = 0; fi < 3; ++fi)
#pragma unroll
                    for (int fj = 0; fj < 3; ++fj) cf[fi * 3 + fj] += kb[fi * 24 + fj];
                }
            const int m = n + (dz * g.ny + dy) * g.nx + dx;
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
              for (int fj = 0; fj < 3; ++fj) {
                const double u = U[(size_t)(fj * K + f0 + q) * NP + m];
#pragma unroll
                for (int d = 0; d < 3; ++d) acc[d][q] += cf[d * 3 + fj] * u;
              }
          }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const float kv = kinv[(size_t)d * NP + n];
      float rf[4], pp[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const size_t idx = (size_t)(d * K + f0 + q) * NP + n;
        const double r64 = (F ? F[idx] : 0.0) - acc[d][q];
        rf[q] = (float)r64;
        pp[q] = kv * rf[q];
        const double z64 = (double)kv * r64;
        zz[q] += z64 * z64;
        rz[q] += (double)rf[q] * (double)pp[q];
        rr[q] += r64 * r64;
      }
      const size_t o = (size_t)(d * KG + grp) * NP + n;
      R[o] = make_float4(rf[0], rf[1], rf[2], rf[3]);
      X[o] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      P[o] = make_float4(pp[0], pp[1], pp[2], pp[3]);
    }
  }
  double v[12];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    v[r] = zz[r];
    v[4 + r] = rz[r];
    v[8 + r] = rr[r];
  }
  block_sum<12, 256>(v, lds);
  if (threadIdx.x == 0)
#pragma unroll
    for (int t = 0; t < 3; ++t)
#pragma unroll
      for (int q = 0; q < 4; ++q) part[(size_t)(t * K + f0 + q) * gridDim.x + blockIdx.x] = v[t * 4 + q];
}

// one block: outer convergence test and the inner CG start state.  active
// (may be null): fields whose entry is 0 count as converged (they take no
// part in the solve).
template <int K>
__global__ void __launch_bounds__(kRedThreads) k_m_outer(int nparts, const double *part, int first, double rel,
                                                         double abs_tol, double tau_min, int max_inner,
                                                         const int *active, MOuter *o, MState *s0) {
  __shared__ double t[3 * kMaxRhs];
  __shared__ int conv_s[kMaxRhs];
  const double s = rows_sum<3 * K>(part, nparts, nparts);
  if (threadIdx.x < 3 * K) t[threadIdx.x] = s;
  __syncthreads();
  // one thread per field; every read of the old state comes before any write
  const int r = threadIdx.x;
  if (r < K) {
    const double pn = sqrt(t[r]), rn = sqrt(t[2 * K + r]);
    const double target = first ? fmax(abs_tol, rel * pn) : o->target[r];
    const int conv = pn < target || pn == 0.0 || (active && !active[r]);
    conv_s[r] = conv;
    if (first) {
      o->pn0[r] = pn;
      o->rn0[r] = rn;
      o->target[r] = target;
    }
    o->pn[r] = pn;
    o->rn[r] = rn;
    o->conv[r] = conv;
    s0->rz[r] = t[K + r];
    s0->pn[r] = pn;
    s0->pn0[r] = pn;
    // aim a little below what the outer test needs, but not below what FP32 can resolve
    s0->tau[r] = (pn > 0.0) ? fmax(tau_min, kInnerAim * target / pn) : 1.0;
    s0->alpha[r] = 0.0;
    s0->beta[r] = 0.0;
    s0->done[r] = conv;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int all = 1;
    for (int q = 0; q < K; ++q) all &= conv_s[q];
    o->all_conv = all;
    s0->its = 0;
    s0->max_its = max_inner;
    s0->all_done = all;
  }
}

// ---------------------------------------------------------------------------
// Inner CG (the reference's Jacobi-PCG recurrence, src/ell.cpp:80-117), one
// iteration = k_m_spmv, k_m_upd, k_m_pupd.
//
// k_m_spmv: AP = A P for every field and the partials of (p, Ap).  One block
// per slice of 64 nodes; wave g handles field group g (fields 4g .. 4g+3), so
// a block has K/4 waves.  Partial table part1[field * nsl + slice].
//
// The slice's matrix streams HBM -> LDS by direct buffer loads (16 bytes per
// lane, no VGPRs held) one stencil row (7 quads, 3 neighbours) at a time,
// two rows ahead in a 3-stage ring; the block's waves share each row (each
// wave moves a third of it), so the matrix crosses L1 once per block instead
// of once per wave.  For each row (dz, dy) a wave gathers p of the dx = 0
// neighbours only, 64 consecutive nodes; the dx = -1 and dx = +1 values are
// those of the adjacent lanes, moved by DPP wave shifts, with lanes 0 and 63
// taking theirs from a second load.  That is a third of the gathers of
// loading every neighbour.
//
// Neighbour addresses are n + offset clamped into [0, NP): every neighbour
// of an interior node lies in the grid, and the rows of boundary and padding
// nodes are zero (with P finite everywhere), so no in-grid test is needed.
// ---------------------------------------------------------------------------
typedef __attribute__((address_space(3))) void lds_void;
constexpr int kStages = 3;
// Buffer resource word 3 for raw 32-bit-element buffers (DATA_FORMAT 32).
constexpr int kBufWord3 = 0x00020000;
static_assert(kRowQuads * 4 == 28 && kRows == 9, "stencil rows of 27 floats + 1 pad");

// s_waitcnt immediate for vmcnt <= n (gfx9 encoding: vmcnt in bits [3:0]
// and [15:14]; expcnt and lgkmcnt at their maxima, i.e. not waited for)
constexpr int vm_wait(int n) { return (n & 15) | ((n >> 4) << 14) | 0x70 | 0xF00; }

// The matrix part of stencil row w is in LDS (this wave's loads) once at
// most `after` vector loads younger than them are outstanding; vector-memory
// loads complete in issue order.
template <int after>
__device__ __forceinline__ void wait_loads() {
  static_assert(after >= 0 && after < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt(vm_wait(after));
}

// Reads the 7 quads of one stage (this lane's column) with one asm block
// that ends with its own lgkmcnt(0).  The compiler cannot tell which stage a
// direct load wrote, so before an LDS read of its own it waits for every
// outstanding vector load, the P gathers in flight included; here the waits
// for the direct loads are the explicit wait_loads + barrier.  Volatile with a
// memory clobber, the block stays between the barrier and the next loads.
__device__ __forceinline__ void read_stage(const float4 *stage, float a[28]) {
  static_assert(kRowQuads == 7, "the asm block reads 7 quads 1024 bytes apart");
  const unsigned la = (unsigned)(size_t)(lds_void *)stage;
  f4v m0, m1, m2, m3, m4, m5, m6;
  asm volatile(
      "ds_read_b128 %0, %7\n\t"
      "ds_read_b128 %1, %7 offset:1024\n\t"
      "ds_read_b128 %2, %7 offset:2048\n\t"
      "ds_read_b128 %3, %7 offset:3072\n\t"
      "ds_read_b128 %4, %7 offset:4096\n\t"
      "ds_read_b128 %5, %7 offset:5120\n\t"
      "ds_read_b128 %6, %7 offset:6144\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(m0), "=&v"(m1), "=&v"(m2), "=&v"(m3), "=&v"(m4), "=&v"(m5), "=&v"(m6)
      : "v"(la)
      : "memory");
  const f4v m[7] = {m0, m1, m2, m3, m4, m5, m6};
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    a[q * 4 + 0] = m[q].x;
    a[q * 4 + 1] = m[q].y;
    a[q * 4 + 2] = m[q].z;
    a[q * 4 + 3] = m[q].w;
  }
}

// DPP wave shifts: lane i takes lane i - 1's value (lane 0 keeps `old`), or
// lane i + 1's (lane 63 keeps `old`)
__device__ __forceinline__ float4 from_left(const float4 &v, const float4 &old) {
  auto f = [](float x, float o) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(o), __float_as_int(x), 0x138, 0xf, 0xf, false));
  };
  return make_float4(f(v.x, old.x), f(v.y, old.y), f(v.z, old.z), f(v.w, old.w));
}
__device__ __forceinline__ float4 from_right(const float4 &v, const float4 &old) {
  auto f = [](float x, float o) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(o), __float_as_int(x), 0x130, 0xf, 0xf, false));
  };
  return make_float4(f(v.x, old.x), f(v.y, old.y), f(v.z, old.z), f(v.w, old.w));
}

template <int K>
__global__ void __launch_bounds__(256) k_m_spmv(Grid g, int NP, int nsl, int per, const float4 *A, const float4 *P,
                                                float4 *AP, double *part1, const MState *cur) {
  constexpr int KG = K / 4;
  constexpr int kDma = (kRowQuads + KG - 1) / KG;  // matrix loads per wave per stencil row
  __shared__ float4 stg[kStages][kRowQuads][64];
  static_assert(sizeof(stg[0][0]) == 1024, "one quad of a 64-node slice");
  if (cur->all_done) return;  // whole grid
  const int slice = xcd_block(per);
  if (slice >= nsl) return;   // whole block
  const int grp = threadIdx.x >> 6, lane = threadIdx.x & 63, r0 = grp * 4;
  const int n = slice * 64 + lane, nxny = g.nx * g.ny;
  const float4 *ablk = A + (size_t)slice * kNodeQuads * 64;
  const float4 *pg = P + (size_t)grp * NP;
  // The slice's matrix as a buffer (raw, 16-byte loads straight to LDS,
  // bounds-checked).  Buffer (MUBUF) rather than global (FLAT) direct loads:
  // the compiler treats an outstanding FLAT load that writes LDS as possibly
  // out of order and waits for every vector load before using any gathered p.
  const __amdgpu_buffer_rsrc_t arsrc =
      __builtin_amdgcn_make_buffer_rsrc((void *)ablk, 0, kNodeQuads * 64 * 16, kBufWord3);
  // Every wave issues kDma loads per row so that the number of loads in
  // flight is known at compile time; quads past the row (q >= 7) repeat the
  // row's last quad into the same LDS slot (the same bytes).
  auto issue = [&](int w) {
#pragma unroll
    for (int u = 0; u < kDma; ++u) {
      const int q = min(grp + KG * u, kRowQuads - 1);
      __builtin_amdgcn_raw_ptr_buffer_load_lds(arsrc, (lds_void *)&stg[w % kStages][q][0], 16,
                                               ((w * kRowQuads + q) * 64 + lane) * 16, 0, 0, 0);
    }
  };
  // A wave whose four fields have stopped still moves its share of the
  // matrix and meets every barrier, but gathers and computes nothing.
  if (cur->done[r0] && cur->done[r0 + 1] && cur->done[r0 + 2] && cur->done[r0 + 3]) {
    issue(0);
    issue(1);
#pragma unroll
    for (int w = 0; w < kRows; ++w) {
      __builtin_amdgcn_s_waitcnt(vm_wait(0));
      __builtin_amdgcn_s_barrier();
      if (w + 2 < kRows) issue(w + 2);
    }
    return;
  }
  float acc[3][4];
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[d][r] = 0.0f;
  // lanes 0 and 63 also load their outer x-neighbour; the other lanes load
  // their own node again (an L1 hit), so the load needs no branch
  const int dedge = lane == 0 ? -1 : (lane == 63 ? 1 : 0);
  issue(0);
  asm volatile("" ::: "memory");  // no gather of row 0 ahead of row 0's matrix loads
  issue(1);
  // Row w: gather p at the dx = 0 neighbours and the edge values (6 loads),
  // wait for the matrix part of row w (the loads of row w + 1 and the
  // gathers issued after it may stay in flight), barrier, read the stage,
  // start row w + 2 into the stage row w - 1 used (every wave has read it
  // before this barrier), then the FMAs of neighbours dx = -1, 0, 1 in the
  // reference order.  The wait counts hold because no load moves across the
  // asm blocks: whatever is younger than row w's matrix loads includes row
  // w + 1's loads and row w's gathers.
#pragma unroll
  for (int w = 0; w < kRows; ++w) {
    const int off = (w / 3 - 1) * nxny + (w % 3 - 1) * g.nx;
    const int m = min(max(n + off, 0), NP - 1), me = min(max(n + off + dedge, 0), NP - 1);
    float4 pm[3], pe[3];
#pragma unroll
    for (int fj = 0; fj < 3; ++fj) pm[fj] = pg[(size_t)fj * KG * NP + m];
#pragma unroll
    for (int fj = 0; fj < 3; ++fj) pe[fj] = pg[(size_t)fj * KG * NP + me];
    asm volatile("" ::: "memory");  // every gather of row w is issued before the wait
    if (w + 1 < kRows)
      wait_loads<6 + kDma>();
    else
      wait_loads<6>();
    __builtin_amdgcn_s_barrier();  // every wave's share of row w is in LDS
    float a[28];
    read_stage(&stg[w % kStages][0][lane], a);
    if (w + 2 < kRows) issue(w + 2);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      float4 v[3];
#pragma unroll
      for (int fj = 0; fj < 3; ++fj) {
        if (j == 0)
          v[fj] = from_left(pm[fj], pe[fj]);
        else if (j == 1)
          v[fj] = pm[fj];
        else
          v[fj] = from_right(pm[fj], pe[fj]);
      }
#pragma unroll
      for (int fj = 0; fj < 3; ++fj)
#pragma unroll
        for (int fi = 0; fi < 3; ++fi) {
          const float c = a[j * 9 + fi * 3 + fj];
          acc[fi][0] += c * v[fj].x;
          acc[fi][1] += c * v[fj].y;
          acc[fi][2] += c * v[fj].z;
          acc[fi][3] += c * v[fj].w;
        }
    }
  }
  // p at the node itself is read again (an L2 hit) rather than kept in
  // registers from row 4
  double pap[4] = {0, 0, 0, 0};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float4 pc = pg[(size_t)d * KG * NP + n];
    AP[(size_t)(d * KG + grp) * NP + n] = make_float4(acc[d][0], acc[d][1], acc[d][2], acc[d][3]);
    pap[0] += (double)pc.x * acc[d][0];
    pap[1] += (double)pc.y * acc[d][1];
    pap[2] += (double)pc.z * acc[d][2];
    pap[3] += (double)pc.w * acc[d][3];
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const double s = wave_sum_dpp(pap[r]);
    if (lane == 0) part1[(size_t)(r0 + r) * nsl + slice] = s;
  }
}

// k_m_upd: alpha = (r, z) / (p, Ap); R -= alpha AP; partials of (z, z) and
// (r, z) with z = kinv R.  Every block re-sums the (p, Ap) partials of its
// four fields (part1[field * nsl + slice]) in the same fixed order, so all
// blocks agree bit for bit; block (0, g) records alpha of group g in the
// state for k_m_pupd.  Grid (blocks of kVecNodes nodes, 3 K/4 components
// c = d * K/4 + g), each thread kVecNodes / 256 nodes; partial tables pz / pr
// [(field * 3 + d) * gridDim.x + block] (the 3 gridDim.x partials of a field
// contiguous; pr = pz + 3 K gridDim.x).  X is advanced in k_m_pupd, which
// reads P anyway.
// COARSE (two-level preconditioner): the block holds whole x-lines, lpb of
// them (lpb nx <= kVecNodes nodes), so that it can also restrict its new r
// along x from LDS: T1 as k_c_rx writes it, the same sums in the same order.
// Its partials then number ceil(ny nz / lpb) per (field, d) (gridDim.x).
template <int K, bool COARSE>
__global__ void __launch_bounds__(kRedThreads) k_m_upd(int NP, int nsl, const double *part1, const float *kinv,
                                                       const float4 *AP, float4 *R, double *pz, double *pr,
                                                       MState *cur, Grid g, CGrid cg, int lpb, float4 *T1) {
  __shared__ double lds[4 * 8];
  __shared__ double al_s[4];
  __shared__ float4 rl[COARSE ? kVecNodes : 1];
  __shared__ float wt[COARSE ? kCoarseMaxLine : 1];
  constexpr int KG = K / 4;
  const int c = blockIdx.y, grp = c % KG, d = c / KG, r0 = grp * 4;
  if (cur->all_done) return;
  if (cur->done[r0] && cur->done[r0 + 1] && cur->done[r0 + 2] && cur->done[r0 + 3]) return;  // whole block
  // the x-hat weight at offset s of a coarse node's support, the same for
  // every coarse node: hat(i0 + s, cx, hx) = hat(1 + s, 1, hx) (published by
  // the barriers below)
  if constexpr (COARSE)
    for (int s2 = threadIdx.x; s2 < 2 * cg.hx - 1; s2 += kRedThreads) wt[s2] = hat(1 + s2, 1, cg.hx);
  constexpr int J = kVecNodes / 256;
  int nbeg = blockIdx.x * kVecNodes, nend = min(nbeg + kVecNodes, NP), l0 = 0, l1 = 0;
  if constexpr (COARSE) {
    l0 = blockIdx.x * lpb;
    l1 = min(l0 + lpb, g.ny * g.nz);
    nbeg = l0 * g.nx;
    nend = l1 * g.nx;
  }
  // The four fields' (p, Ap) partials (one per SpMV slice) are loaded first,
  // then the block's vectors, which stay in flight while the partials are
  // summed and alpha is formed.
  double sp[4];
  rows_load<4, 8>(part1 + (size_t)r0 * nsl, nsl, nsl, sp);
  asm volatile("" ::: "memory");
  float kvv[J];
  float4 apv[J], rvv[J];
#pragma unroll
  for (int jn = 0; jn < J; ++jn) {
    const int nl = min(nbeg + (int)threadIdx.x + jn * 256, nend - 1);  // clamped; the update skips n >= nend
    const size_t o = (size_t)c * NP + nl;
    kvv[jn] = kinv[(size_t)d * NP + nl];
    apv[jn] = AP[o];
    rvv[jn] = R[o];
  }
  const double pap = rows_finish<4>(sp);
  if (threadIdx.x < 4) {
    const int r = r0 + threadIdx.x;
    const double a = cur->done[r] ? 0.0 : cur->rz[r] / pap;
    al_s[threadIdx.x] = a;
    if (blockIdx.x == 0 && d == 0 && !cur->done[r]) cur->alpha[r] = a;
  }
  __syncthreads();
  const float4 al = make_float4((float)al_s[0], (float)al_s[1], (float)al_s[2], (float)al_s[3]);
  double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int jn = 0; jn < J; ++jn) {
    const int n = nbeg + threadIdx.x + jn * 256;
    if (n >= nend) break;
    const size_t o = (size_t)c * NP + n;
    const float kv = kvv[jn];
    const float4 ap = apv[jn];
    float4 r = rvv[jn];
    r.x -= al.x * ap.x;
    r.y -= al.y * ap.y;
    r.z -= al.z * ap.z;
    r.w -= al.w * ap.w;
    R[o] = r;
    if constexpr (COARSE) rl[n - nbeg] = r;
    const float rv[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float z = kv * rv[q];
      v[q] += (double)z * z;
      v[4 + q] += (double)rv[q] * z;
    }
  }
  // block sums: DPP within each wave, then the 4 waves in order
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const double w = wave_sum_dpp(v[q]);
    if ((threadIdx.x & 63) == 0) lds[(threadIdx.x >> 6) * 8 + q] = w;
  }
  __syncthreads();
  if (threadIdx.x < 8) {
    const int q = threadIdx.x;
    const double t = lds[q] + lds[8 + q] + lds[16 + q] + lds[24 + q];
    const size_t o = ((size_t)(r0 + (q & 3)) * 3 + d) * gridDim.x + blockIdx.x;
    if (q < 4)
      pz[o] = t;
    else if (!COARSE)  // the coarse variant's (r, z) partials come from k_c_zdot
      pr[o] = t;
  }
  if constexpr (COARSE) {
    // restriction along x of the block's lines (interior lines only)
    for (int idx = threadIdx.x; idx < (l1 - l0) * cg.mx; idx += kRedThreads) {
      const int line = idx / cg.mx, cx = idx - line * cg.mx + 1, L = l0 + line;
      const int k = L / g.ny, j = L - k * g.ny;
      if (j == 0 || k == 0 || j == g.ny - 1 || k == g.nz - 1) continue;
      const float4 *rline = rl + line * g.nx + (cx - 1) * cg.hx + 1;
      const int n = 2 * cg.hx - 1;
      float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll 3
      for (int s2 = 0; s2 < n; ++s2) {
        const float w = wt[s2];
        const float4 rv = rline[s2];
        acc.x += w * rv.x;
        acc.y += w * rv.y;
        acc.z += w * rv.z;
        acc.w += w * rv.w;
      }
      T1[((size_t)(c * g.nz + k) * g.ny + j) * cg.mx + (cx - 1)] = acc;
    }
  }
}

// The reference's CG step after the residual update (src/ell.cpp:95-116) for
// field r: ||z|| from (z, z), the stopping test, beta = (r, z)_new / (r, z)_old.
struct MStep {
  double pn, beta, rz;
  int done;
};
__device__ __forceinline__ MStep m_step(const MState *cur, int r, double zz, double rz) {
  MStep o;
  const int was = cur->done[r];
  o.pn = was ? cur->pn[r] : sqrt(zz);
  o.done = was || o.pn < cur->tau[r] * cur->pn0[r] || !(o.pn == o.pn);
  o.rz = was ? cur->rz[r] : rz;
  o.beta = was ? 0.0 : rz / cur->rz[r];
  return o;
}

// ---------------------------------------------------------------------------
// Two-level preconditioner: z = D^-1 r + P_c A_c^-1 P_c^T r, with P_c the
// trilinear hats of the interior nodes of a coarse grid aligned with the fine
// one (cg.hx fine elements per coarse element along x, ...), A_c = P_c^T A
// P_c (Galerkin; assembled and inverted once on the device, coarse.hip, and
// kept as a dense FP32 inverse).  Per application: restriction along x (k_c_rx), then along y and
// z (k_c_ryz), the dense coarse solve in kCoarseSplit column splits
// (k_c_gemv) and their sum (k_c_gemv_sum), then z with the prolongation
// (k_c_zdot), which k_m_pupd reads.  Restriction and prolongation use the
// same hat weights (their products are rounded in different orders, a
// difference at FP32 rounding level), so the preconditioner is symmetric to
// that level.  All sums run in fixed orders.  Every thread's loads are
// independent of each other, so they are in flight together: these kernels
// are short and latency decides their time.
// ---------------------------------------------------------------------------

// One fine index along one side: its two coarse nodes (0-based interior
// indices, clamped) and their hat weights (hat_pair)
struct AxisW {
  int q[2];
  float w[2];
};

// The three sides' tables of a block (x: nx entries, then y, then z), filled
// by the whole block; every thread must reach the barrier.
__device__ __forceinline__ void axis_tables(const Grid &g, const CGrid &cg, AxisW (*tab)[kCoarseMaxLine]) {
  for (int t = threadIdx.x; t < g.nx + g.ny + g.nz; t += blockDim.x) {
    const int a = t < g.nx ? 0 : t < g.nx + g.ny ? 1 : 2;
    const int i = a == 0 ? t : a == 1 ? t - g.nx : t - g.nx - g.ny;
    const int h = a == 0 ? cg.hx : a == 1 ? cg.hy : cg.hz, m = a == 0 ? cg.mx : a == 1 ? cg.my : cg.mz;
    AxisW e;
    hat_pair(i, h, m, e.w, e.q);
    tab[a][i] = e;
  }
  __syncthreads();
}

// P_c e at a fine node whose sides' table entries are ax, ay, az, from the
// block's staged coarse corrections el (coarse z-levels czlo .. czhi, every
// y and x, of the block's direction and field group): the 8 coarse nodes
// around the node, all read at once.  A corner outside the grid has weight 0
// and index 0, which may lie below czlo: its level is clamped into the staged
// ones.
__device__ __forceinline__ float4 prolong(const CGrid &cg, const float4 *el, int czlo, int czhi, const AxisW &ax,
                                          const AxisW &ay, const AxisW &az) {
  float4 e[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const int cz = min(max(az.q[t >> 2], czlo), czhi) - czlo;
    e[t] = el[(cz * cg.my + ay.q[(t >> 1) & 1]) * cg.mx + ax.q[t & 1]];
  }
  float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const float w = ax.w[t & 1] * (ay.w[(t >> 1) & 1] * az.w[t >> 2]);
    acc.x += w * e[t].x;
    acc.y += w * e[t].y;
    acc.z += w * e[t].z;
    acc.w += w * e[t].w;
  }
  return acc;
}

// Restriction along x: one thread per (component c = d KG + grp, interior
// x-line (j, k), interior coarse x node cx): T1[((c nz + k) ny + j) mx + cx -
// 1] = sum_i hat_cx(i) R(i, j, k) over the 2 hx - 1 nodes where the hat is
// non-zero.  Lines on the y or z boundary are not needed (the y / z hats
// vanish there).
template <int K>
__global__ void __launch_bounds__(256) k_c_rx(Grid g, CGrid cg, int NP, const float4 *R, float4 *T1,
                                              const MState *cur) {
  constexpr int KG = K / 4;
  if (cur->all_done) return;
  const int ly = g.ny - 2, lz = g.nz - 2;
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= 3 * KG * ly * lz * cg.mx) return;
  const int cx = t % cg.mx + 1;
  int rest = t / cg.mx;
  const int j = rest % ly + 1;
  rest /= ly;
  const int k = rest % lz + 1, c = rest / lz;
  const float4 *line = R + (size_t)c * NP + ((size_t)k * g.ny + j) * g.nx;
  const int i0 = (cx - 1) * cg.hx + 1, n = 2 * cg.hx - 1;
  float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll 8
  for (int s = 0; s < n; ++s) {
    const float w = hat(i0 + s, cx, cg.hx);
    const float4 v = line[i0 + s];
    acc.x += w * v.x;
    acc.y += w * v.y;
    acc.z += w * v.z;
    acc.w += w * v.w;
  }
  T1[((size_t)(c * g.nz + k) * g.ny + j) * cg.mx + (cx - 1)] = acc;
}

// Restriction along y and z: one block per (interior coarse y-z pair, component
// c), for all mx coarse x nodes at once.  Thread t takes coarse x node t % mx
// and the (j, k) lines slot, slot + nslot, ... of the pair's window
// ((2 hy - 1)(2 hz - 1) lines, slot = t / mx, nslot = 256 / mx), so a line's
// mx x-restricted values are read together; the slots' sums are added in LDS
// in slot order.  (The window lines of neighbouring pairs overlap: each value
// is read by up to 4 blocks.)
template <int K>
__global__ void __launch_bounds__(256) k_c_ryz(Grid g, CGrid cg, const float4 *T1, float4 *rc4, const MState *cur) {
  constexpr int KG = K / 4;
  __shared__ float4 part[256];
  if (cur->all_done) return;
  const int c = blockIdx.y, d = c / KG, grp = c % KG;
  const int cy = blockIdx.x % cg.my + 1, cz = blockIdx.x / cg.my + 1;
  const int mx = cg.mx, nslot = 256 / mx, t = threadIdx.x;
  const int wy = 2 * cg.hy - 1, npair = wy * (2 * cg.hz - 1);
  const int j0 = (cy - 1) * cg.hy + 1, k0 = (cz - 1) * cg.hz + 1;
  float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (t < nslot * mx) {
    const int cx = t % mx;
    for (int s = t / mx; s < npair; s += nslot) {
      const int j = j0 + s % wy, k = k0 + s / wy;
      const float w = hat(j, cy, cg.hy) * hat(k, cz, cg.hz);
      const float4 v = T1[((size_t)(c * g.nz + k) * g.ny + j) * mx + cx];
      acc.x += w * v.x;
      acc.y += w * v.y;
      acc.z += w * v.z;
      acc.w += w * v.w;
    }
  }
  part[t] = acc;
  __syncthreads();
  if (t < mx) {
    float4 sum = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (int sl = 0; sl < nslot; ++sl) {
      const float4 v = part[sl * mx + t];
      sum.x += v.x;
      sum.y += v.y;
      sum.z += v.z;
      sum.w += v.w;
    }
    const int q = ((cz - 1) * cg.my + (cy - 1)) * mx + t;
    rc4[(size_t)(q * 3 + d) * KG + grp] = sum;
  }
}

// The coarse solve e = A_c^-1 r_c (nc x nc times nc x K), split over the
// columns: block (row tile of 64, split s) sums the columns of split s, each
// wave a quarter of them (lane = row; the symmetric inverse is read by column,
// one coalesced line per column), the waves combined in order, into ecp;
// k_c_gemv_sum adds the splits.  (Summing them in the last block to finish,
// counted with an atomic, measured 9x slower: the device-scope fences it needs
// write the L2 back.)
template <int K>
__global__ void __launch_bounds__(256) k_c_gemv(CGrid cg, const float *__restrict__ ainv,
                                                const float *__restrict__ rc, float *__restrict__ ecp,
                                                const MState *cur) {
  __shared__ float red[3][64][K];
  if (cur->all_done) return;
  const int nc = cg.nc, lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int row = blockIdx.x * 64 + lane, split = blockIdx.y;
  const int chunk = (nc + kCoarseSplit - 1) / kCoarseSplit;
  const int c0 = split * chunk, c1 = min(nc, c0 + chunk);
  const int rowc = min(row, nc - 1);  // loads of rows past nc are clamped, their sums not stored
  float acc[K];
#pragma unroll
  for (int f = 0; f < K; ++f) acc[f] = 0.0f;
#pragma unroll 16
  for (int col = c0 + w; col < c1; col += 4) {
    const float a = ainv[(size_t)col * nc + rowc];
    const float *x = rc + (size_t)col * K;
#pragma unroll
    for (int f = 0; f < K; ++f) acc[f] += a * x[f];
  }
  if (w > 0)
#pragma unroll
    for (int f = 0; f < K; ++f) red[w - 1][lane][f] = acc[f];
  __syncthreads();
  if (w == 0 && row < nc)
#pragma unroll
    for (int f = 0; f < K; ++f)
      ecp[((size_t)split * nc + row) * K + f] = acc[f] + red[0][lane][f] + red[1][lane][f] + red[2][lane][f];
}

template <int K>
__global__ void __launch_bounds__(256) k_c_gemv_sum(CGrid cg, const float *ecp, float *ec, const MState *cur) {
  if (cur->all_done) return;
  const int t = blockIdx.x * 256 + threadIdx.x;
  const size_t n = (size_t)cg.nc * K;
  if (t >= (int)n) return;
  float v[kCoarseSplit];
#pragma unroll
  for (int p = 0; p < kCoarseSplit; ++p) v[p] = ecp[(size_t)p * n + t];
  float s = 0.0f;
#pragma unroll
  for (int p = 0; p < kCoarseSplit; ++p) s += v[p];
  ec[t] = s;
}

// z = D^-1 r + P_c e into Z (START: into P, the inner CG's first direction)
// and the partials of (r, z) in k_m_upd's layout pr[(field * 3 + d) *
// gridDim.x + block] (grid as k_m_upd: kVecNodes nodes per block, here on
// kZdotThreads threads, so that each thread's rounds of dependent gathers are
// few).  Loads are clamped to valid nodes (their contributions dropped).
constexpr int kZdotThreads = 512;
template <int K, bool START>
__global__ void __launch_bounds__(kZdotThreads) k_c_zdot(Grid g, CGrid cg, int NP, const float *kinv,
                                                         const float4 *R, const float4 *ec4, float4 *Z, double *pr,
                                                         const MState *cur) {
  constexpr int kW = kZdotThreads / 64;
  __shared__ double lds[kW * 4];
  constexpr int KG = K / 4;
  const int c = blockIdx.y, grp = c % KG, d = c / KG, r0 = grp * 4;
  if (!START) {
    if (cur->all_done) return;
    if (cur->done[r0] && cur->done[r0 + 1] && cur->done[r0 + 2] && cur->done[r0 + 3]) return;  // whole block
  }
  __shared__ AxisW tab[3][kCoarseMaxLine];
  __shared__ float4 el[kCoarseMaxDofs / 3 + 1];
  // the coarse corrections of this block's nodes, staged: coarse z-levels
  // k / hz - 1 and k / hz (0-based, where inside 0 .. mz - 1) of its planes
  // (a block of the last plane alone, k = nz - 1, where every hat is zero,
  // stages one level)
  {
    const int nxy = g.nx * g.ny, nlo = blockIdx.x * kVecNodes, nhi = min(nlo + kVecNodes, NP) - 1;
    const int zhi = min(cg.mz - 1, min(nhi / nxy, g.nz - 1) / cg.hz);
    const int zlo = min(zhi, max(0, min(nlo / nxy, g.nz - 1) / cg.hz - 1));
    const int plane = cg.mx * cg.my;
    for (int t = threadIdx.x; t < (zhi - zlo + 1) * plane; t += kZdotThreads)
      el[t] = ec4[(size_t)((zlo * plane + t) * 3 + d) * KG + grp];
  }
  axis_tables(g, cg, tab);  // (its barrier also publishes el)
  const int nxy0 = g.nx * g.ny, n0lo = blockIdx.x * kVecNodes, n0hi = min(n0lo + kVecNodes, NP) - 1;
  const int czhi = min(cg.mz - 1, min(n0hi / nxy0, g.nz - 1) / cg.hz);
  const int czlo = min(czhi, max(0, min(n0lo / nxy0, g.nz - 1) / cg.hz - 1));
  constexpr int J = kVecNodes / kZdotThreads;
  const int n0 = blockIdx.x * kVecNodes + threadIdx.x;
  // the node's (i, j, k), stepped by kZdotThreads nodes per round (nodes past
  // the grid, k >= nz, read the last plane's entries; their part is dropped)
  int i = n0 % g.nx, j = (n0 / g.nx) % g.ny, k = n0 / (g.nx * g.ny);
  const int di = kZdotThreads % g.nx, dl = kZdotThreads / g.nx;
  double v[4] = {0, 0, 0, 0};
#pragma unroll
  for (int jn = 0; jn < J; ++jn) {
    const int n = n0 + jn * kZdotThreads;
    const int nl = min(n, NP - 1);
    const size_t o = (size_t)c * NP + nl;
    const float kv = kinv[(size_t)d * NP + nl];
    const float4 r = R[o];
    const float4 e = prolong(cg, el, czlo, czhi, tab[0][i], tab[1][j], tab[2][min(k, g.nz - 1)]);
    const bool real = n < g.nn;  // padding nodes have r = 0 and no coarse part
    i += di;
    int carry = dl;
    if (i >= g.nx) {
      i -= g.nx;
      ++carry;
    }
    j += carry;
    while (j >= g.ny) {
      j -= g.ny;
      ++k;
    }
    const float4 z = make_float4(kv * r.x + (real ? e.x : 0.0f), kv * r.y + (real ? e.y : 0.0f),
                                 kv * r.z + (real ? e.z : 0.0f), kv * r.w + (real ? e.w : 0.0f));
    if (n < NP) {
      Z[o] = z;
      v[0] += (double)r.x * z.x;
      v[1] += (double)r.y * z.y;
      v[2] += (double)r.z * z.z;
      v[3] += (double)r.w * z.w;
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const double w = wave_sum_dpp(v[q]);
    if ((threadIdx.x & 63) == 0) lds[(threadIdx.x >> 6) * 4 + q] = w;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    const int q = threadIdx.x;
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kW; ++w) t += lds[w * 4 + q];
    pr[((size_t)(r0 + q) * 3 + d) * gridDim.x + blockIdx.x] = t;
  }
}

// start: the state's (r, z) of every field from k_c_zdot<K, true>'s partials
template <int K>
__global__ void __launch_bounds__(kRedThreads) k_c_rzfix(int nbx, const double *prz, MState *s0) {
  const size_t row = (size_t)3 * nbx;
  const double rz = rows_sum<K>(prz, row, (int)row);
  if (threadIdx.x < K) s0->rz[threadIdx.x] = rz;
}

// k_m_pupd: X += alpha P; P = kinv R + beta P for the fields still running;
// the next CG state.  Every block re-sums the (z, z) and (r, z) partials of
// its four fields in the same fixed order (so all blocks agree bit for bit),
// and block (0, 0) sums all fields and writes the state of the next iteration
// (beta, ||z||, stop flags, iteration count).  Grid as k_m_upd.
// COARSE (two-level preconditioner): z is read from Z (k_c_zdot) instead of
// being formed from kinv and R.
template <int K, bool COARSE>
__global__ void __launch_bounds__(kRedThreads) k_m_pupd(int NP, const float *kinv, const float4 *R, const float4 *Z,
                                                        float4 *P, float4 *X, const double *pz, int nbz,
                                                        const double *pr, const MState *cur,
                                                        MState *nxt) {
  constexpr int KG = K / 4;
  __shared__ double beta_s[kMaxRhs];
  __shared__ int live_s[kMaxRhs], dn_s[kMaxRhs];
  const int c = blockIdx.y, grp = c % KG, d = c / KG, r0 = grp * 4;
  const bool head = blockIdx.x == 0 && blockIdx.y == 0;
  if (cur->all_done) {
    if (head) {  // keep the finished state for the iterations queued after it
      const int *s = reinterpret_cast<const int *>(cur);
      int *t = reinterpret_cast<int *>(nxt);
      for (int q = threadIdx.x; q < (int)(sizeof(MState) / sizeof(int)); q += kRedThreads) t[q] = s[q];
    }
    return;
  }
  const size_t row = (size_t)3 * gridDim.x, rowz = (size_t)3 * nbz;  // (r, z) / (z, z) partials per field
  // The block's (z, z) / (r, z) partials are loaded first (block (0, 0) sums
  // every field's after its vectors), then this block's vectors (rows
  // clamped; the update skips n >= NP), which stay in flight while the
  // partials are summed.
  double szz[4] = {}, srz[4] = {};
  if (!head) {
    rows_load<4>(pz + r0 * rowz, rowz, (int)rowz, szz);
    rows_load<4>(pr + r0 * row, row, (int)row, srz);
  }
  asm volatile("" ::: "memory");
  constexpr int J = kVecNodes / 256;
  float kvv[J];
  float4 rvv[J], pvv[J], xvv[J];
  const int n0 = blockIdx.x * kVecNodes + threadIdx.x;
#pragma unroll
  for (int jn = 0; jn < J; ++jn) {
    const int n = min(n0 + jn * 256, NP - 1);
    const size_t o = (size_t)c * NP + n;
    if constexpr (COARSE) {
      kvv[jn] = 0.0f;
      rvv[jn] = Z[o];  // z itself
    } else {
      kvv[jn] = kinv[(size_t)d * NP + n];
      rvv[jn] = R[o];
    }
    pvv[jn] = P[o];
    xvv[jn] = X[o];
  }
  const int its = cur->its + 1;
  if (head) {
    const double zz = rows_sum<K>(pz, rowz, (int)rowz);
    const double rz = rows_sum<K>(pr, row, (int)row);
    const int r = threadIdx.x;
    if (r < K) {
      const MStep o = m_step(cur, r, zz, rz);
      nxt->rz[r] = o.rz;
      nxt->beta[r] = o.beta;
      nxt->alpha[r] = cur->alpha[r];
      nxt->pn[r] = o.pn;
      nxt->pn0[r] = cur->pn0[r];
      nxt->tau[r] = cur->tau[r];
      nxt->done[r] = o.done;
      dn_s[r] = o.done;
      if (r >= r0 && r < r0 + 4) {
        beta_s[r - r0] = o.beta;
        live_s[r - r0] = !o.done && its < cur->max_its;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int all = 1;
      for (int q = 0; q < K; ++q) all &= dn_s[q];
      nxt->its = its;
      nxt->max_its = cur->max_its;
      nxt->all_done = all || its >= cur->max_its;
    }
  } else {
    const double zz = rows_finish<4>(szz);
    const double rz = rows_finish<4>(srz);
    if (threadIdx.x < 4) {
      const MStep o = m_step(cur, r0 + threadIdx.x, zz, rz);
      beta_s[threadIdx.x] = o.beta;
      live_s[threadIdx.x] = !o.done && its < cur->max_its;
    }
  }
  __syncthreads();
  const bool was[4] = {cur->done[r0] != 0, cur->done[r0 + 1] != 0, cur->done[r0 + 2] != 0, cur->done[r0 + 3] != 0};
  if (was[0] && was[1] && was[2] && was[3]) return;  // the group stopped before this iteration
  const float4 al = make_float4(was[0] ? 0.0f : (float)cur->alpha[r0], was[1] ? 0.0f : (float)cur->alpha[r0 + 1],
                                was[2] ? 0.0f : (float)cur->alpha[r0 + 2], was[3] ? 0.0f : (float)cur->alpha[r0 + 3]);
  const float be[4] = {(float)beta_s[0], (float)beta_s[1], (float)beta_s[2], (float)beta_s[3]};
  const int lv[4] = {live_s[0], live_s[1], live_s[2], live_s[3]};
#pragma unroll
  for (int jn = 0; jn < J; ++jn) {
    const int n = n0 + jn * 256;
    if (n >= NP) break;
    const size_t o = (size_t)c * NP + n;
    const float kv = kvv[jn];
    const float4 r = rvv[jn];
    float4 p = pvv[jn], x = xvv[jn];
    x.x += al.x * p.x;
    x.y += al.y * p.y;
    x.z += al.z * p.z;
    x.w += al.w * p.w;
    X[o] = x;
    const float4 z = COARSE ? r : make_float4(kv * r.x, kv * r.y, kv * r.z, kv * r.w);
    if (lv[0]) p.x = z.x + be[0] * p.x;
    if (lv[1]) p.y = z.y + be[1] * p.y;
    if (lv[2]) p.z = z.z + be[2] * p.z;
    if (lv[3]) p.w = z.w + be[3] * p.w;
    P[o] = p;
  }
}

// U += X (FP64 update of the defect-correction step): grid (node blocks,
// 3 K/4 components); field 4g + q of component (d, g) is U row d * K + 4g + q
template <int K>
__global__ void __launch_bounds__(256) k_m_accum(int NP, const float4 *X, double *U) {
  constexpr int KG = K / 4;
  const int c = blockIdx.y, grp = c % KG, d = c / KG;
  const int n = blockIdx.x * 256 + threadIdx.x;
  if (n >= NP) return;
  const float4 x = X[(size_t)c * NP + n];
  double *u = U + (size_t)(d * K + grp * 4) * NP + n;
  u[0] += (double)x.x;
  u[(size_t)NP] += (double)x.y;
  u[(size_t)2 * NP] += (double)x.z;
  u[(size_t)3 * NP] += (double)x.w;
}

// ---------------------------------------------------------------------------
// calc_ave_stress (src/average.cpp:57-82) of the K fields, FP64, for linear
// elastic phases (the only ones the multi-RHS solver takes): an element's
// Gauss-weighted sum of D B_gp u_e over its 8 points is D (sum_gp wg B_gp) u_e,
// so each element applies the weighted sum of the 8 B matrices once (8x fewer
// products than point by point; the sums are rounded in another order).
// Partials part[(field * 6 + v) * gridDim.x + block], then one block per
// output.
// ---------------------------------------------------------------------------
template <int K>
__global__ void __launch_bounds__(256) k_m_ave(Grid g, int NP, const double *bmat_g, const material_base *mats,
                                               const int *et, const double *U, double *part) {
  __shared__ double bw[144];
  __shared__ double lds[4 * 6];
  for (int t = threadIdx.x; t < 144; t += blockDim.x) {
    double s = 0.0;
    for (int gp = 0; gp < 8; ++gp) s += bmat_g[gp * 144 + t] * g.wg;
    bw[t] = s;
  }
  __syncthreads();
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  int ex = 0, ey = 0, ez = 0, ph = 0;
  const bool act = e < g.nelem;
  if (act) {
    ex = e % g.nex;
    ey = (e / g.nex) % g.ney;
    ez = e / (g.nex * g.ney);
    ph = et[e];
  }
  for (int r = 0; r < K; ++r) {
    double s6[6] = {0, 0, 0, 0, 0, 0};
    if (act) {
      double ue[24];
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        const int node = ((ez + kHexPos[m][2]) * g.ny + (ey + kHexPos[m][1])) * g.nx + ex + kHexPos[m][0];
#pragma unroll
        for (int d = 0; d < 3; ++d) ue[m * 3 + d] = U[(size_t)(d * K + r) * NP + node];
      }
      double eps[6];
#pragma unroll
      for (int v = 0; v < 6; ++v) {
        double s = 0;
#pragma unroll
        for (int q = 0; q < 24; ++q) s += bw[v * 24 + q] * ue[q];
        eps[v] = s;
      }
      mat_stress(mats[ph], eps, nullptr, s6);
    }
    block_sum<6, 256>(s6, lds);
    if (threadIdx.x < 6) part[(size_t)(r * 6 + threadIdx.x) * gridDim.x + blockIdx.x] = s6[threadIdx.x];
  }
}

// ---------------------------------------------------------------------------
// Adjoint derivatives (elastic).  With R(u) = K u - F_bc the interior
// equations and lambda_v the adjoint fields (fields 0..5 of U), for any
// parameter p of the boundary values or of the stiffness
//   d sigma_v / dp = sum_e sum_gp wg (e_v / V - B lambda_v)^T dC/dp-or-C (B du/dp-or-u)
// so both derivatives are single passes over the elements:
//   d sigma_v / d eps_i : C, and w_i = the boundary values of the unit strain
//                         e_i (zero inside), which only boundary elements see;
//   d sigma_v / d E_p, d nu_p : dC/dE, dC/dnu of the phase p of the element,
//                         and the forward field u of the strain.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool elem_touches_boundary(const Grid &g, int ex, int ey, int ez) {
  return ex == 0 || ey == 0 || ez == 0 || ex == g.nex - 1 || ey == g.ney - 1 || ez == g.nez - 1;
}

// boundary value of unit strain i at node (i, j, k): the coordinate rule of
// set_displ_bc (src/micro3D.cpp:26-78)
__device__ __forceinline__ void unit_bc(const Grid &g, int ui, int i, int j, int k, double *w) {
  const double L = 1.0;
  double x[3];
  if (i == g.nx - 1) {
    x[0] = L, x[1] = j * g.dy, x[2] = k * g.dz;
  } else if (i == 0) {
    x[0] = 0, x[1] = j * g.dy, x[2] = k * g.dz;
  } else if (j == g.ny - 1) {
    x[0] = i * g.dx, x[1] = L, x[2] = k * g.dz;
  } else if (j == 0) {
    x[0] = i * g.dx, x[1] = 0, x[2] = k * g.dz;
  } else if (k == g.nz - 1) {
    x[0] = i * g.dx, x[1] = j * g.dy, x[2] = L;
  } else {
    x[0] = i * g.dx, x[1] = j * g.dy, x[2] = 0;
  }
  w[0] = w[1] = w[2] = 0.0;
  if (ui < 3) {
    w[ui] = x[ui];
  } else {
    const int a = ui == 5 ? 1 : 0, b = ui == 3 ? 1 : 2;  // 3: xy, 4: xz, 5: yz
    w[a] = 0.5 * x[b];
    w[b] = 0.5 * x[a];
  }
}

// The elements with a node on the boundary (the outer layer), in a compact
// order: layer ez = 0 whole, the rings of the middle layers, layer nez - 1
// whole.  boundary_elems: how many; boundary_elem: element idx of them.
__host__ __device__ inline int boundary_ring(const Grid &g) {
  const int full = g.nex * g.ney;
  return g.nex > 2 && g.ney > 2 ? full - (g.nex - 2) * (g.ney - 2) : full;
}
__host__ __device__ inline int boundary_elems(const Grid &g) {
  const int full = g.nex * g.ney;
  return g.nez == 1 ? full : 2 * full + (g.nez - 2) * boundary_ring(g);
}
__device__ __forceinline__ void boundary_elem(const Grid &g, int idx, int &ex, int &ey, int &ez) {
  const int full = g.nex * g.ney, ring = boundary_ring(g), mid = g.nez >= 2 ? (g.nez - 2) * ring : 0;
  if (idx < full || idx >= full + mid) {
    const int rem = idx < full ? idx : idx - full - mid;
    ex = rem % g.nex;
    ey = rem / g.nex;
    ez = idx < full ? 0 : g.nez - 1;
    return;
  }
  const int rem = idx - full, r = rem % ring;
  ez = 1 + rem / ring;
  if (ring == full) {
    ex = r % g.nex;
    ey = r / g.nex;
  } else if (r < g.nex) {
    ex = r, ey = 0;
  } else if (r < 2 * g.nex) {
    ex = r - g.nex, ey = g.ney - 1;
  } else if (r < 2 * g.nex + g.ney - 2) {
    ex = 0, ey = 1 + r - 2 * g.nex;
  } else {
    ex = g.nex - 1, ey = 1 + r - 2 * g.nex - (g.ney - 2);
  }
}

// d sigma_v / d eps_i: only elements with a boundary node see w_i.  With the
// element's Gauss sums folded (elastic phases), its term is
//   sum_gp wg ((C B w_i)_v / V - (B lambda_v)^T C B w_i)
//     = (C Bw w_i)_v / V - lambda_v^T Ke w_i,   Bw = sum_gp wg B_gp,
// Ke = sum_gp B^T (C B wg) the phase's element stiffness (ke, the same sums
// as the multi-RHS matrix's).  Grid (boundary-element blocks, 6): blockIdx.y =
// i; a thread forms Ke w_i and C Bw w_i once for the six v.
// part[(v * 6 + i) * gridDim.x + block].
template <int K>
__global__ void __launch_bounds__(256) k_m_adj_ctan(Grid g, int NP, const double *bmat_g, const material_base *mats,
                                                    const int *et, const double *ke, const double *U, double *part) {
  __shared__ double bw[144];
  __shared__ double ke_s[3 * 576];
  __shared__ double lds[4 * 6];
  for (int t = threadIdx.x; t < 144; t += blockDim.x) {
    double s = 0.0;
    for (int gp = 0; gp < 8; ++gp) s += bmat_g[gp * 144 + t] * g.wg;
    bw[t] = s;
  }
  for (int t = threadIdx.x; t < 3 * 576; t += blockDim.x) ke_s[t] = ke[t];
  __syncthreads();
  const int ui = blockIdx.y;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  if (idx < boundary_elems(g)) {
    int ex, ey, ez;
    boundary_elem(g, idx, ex, ey, ez);
    const int e = (ez * g.ney + ey) * g.nex + ex, ph = et[e];
    int node[8];
    double we[24];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const int ni = ex + kHexPos[a][0], nj = ey + kHexPos[a][1], nk = ez + kHexPos[a][2];
      node[a] = (nk * g.ny + nj) * g.nx + ni;
      double w3[3] = {0.0, 0.0, 0.0};  // (a small array of its own: unit_bc indexes it at run time)
      if (node_on_boundary(g, ni, nj, nk)) unit_bc(g, ui, ni, nj, nk, w3);
#pragma unroll
      for (int d = 0; d < 3; ++d) we[a * 3 + d] = w3[d];
    }
    double ew[6], sw[6], t[24];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      double a1 = 0.0;
#pragma unroll
      for (int c = 0; c < 24; ++c) a1 += bw[q * 24 + c] * we[c];
      ew[q] = a1;
    }
    lin_stress(mats[ph], ew, sw);
    const double *kp = ke_s + ph * 576;
#pragma unroll
    for (int r = 0; r < 24; ++r) {
      double a1 = 0.0;
#pragma unroll
      for (int c = 0; c < 24; ++c) a1 += kp[r * 24 + c] * we[c];
      t[r] = a1;
    }
#pragma unroll
    for (int v = 0; v < 6; ++v) {
      double dl = 0.0;
#pragma unroll
      for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int d = 0; d < 3; ++d) dl += U[(size_t)(d * K + v) * NP + node[a]] * t[a * 3 + d];
      acc[v] = sw[v] / g.vol_tot - dl;
    }
  }
  block_sum<6, 256>(acc, lds);
  if (threadIdx.x < 6) part[(size_t)(threadIdx.x * 6 + ui) * gridDim.x + blockIdx.x] = acc[threadIdx.x];
}

// d lambda / d p and d mu / d p of an elastic phase (p = E: which 0, nu: 1),
// as a material whose lin_stress is dC/dp . eps
__device__ __forceinline__ material_base dmat(const material_base &m, int which) {
  material_base d = m;
  const double E = m.E, nu = m.nu;
  if (which == 0) {
    d.lambda = nu / ((1 + nu) * (1 - 2 * nu));
    d.mu = 1 / (2 * (1 + nu));
  } else {
    d.lambda = E * (1 + 2 * nu * nu) / ((1 + nu) * (1 + nu) * (1 - 2 * nu) * (1 - 2 * nu));
    d.mu = -E / (2 * (1 + nu) * (1 + nu));
  }
  return d;
}

// d sigma_v / d theta for theta = E, nu of the element's phase: with the Gauss
// sums folded (elastic phases) an element's term is
//   (dC Bw u_s)_v / V - lambda_v^T dKe u_s,  dKe = sum_gp wg B^T dC B,
// and since C = lambda N + mu M (N: ones on the normal 3x3 block, M =
// diag(2, 2, 2, 1, 1, 1)), dKe = dlambda KN + dmu KM with KN, KM the element
// matrices of N and M (built per block in LDS).  Grid (element blocks, ns):
// blockIdx.y = s, forward field s0 + s; a thread forms dKe u_s once for the six
// v.  part[((s * 6 + v) * 6 + phase * 2 + {E: 0, nu: 1}) * gridDim.x + block].
template <int K>
__global__ void __launch_bounds__(256) k_m_adj_dmat(Grid g, int NP, const double *bmat_g, const material_base *mats,
                                                    const int *et, const double *U, int s0, double *part) {
  __shared__ double bw[144];
  __shared__ double kn[576], km[576];
  __shared__ double lds[4 * 6];
  for (int t = threadIdx.x; t < 144; t += blockDim.x) {
    double sum = 0.0;
    for (int gp = 0; gp < 8; ++gp) sum += bmat_g[gp * 144 + t] * g.wg;
    bw[t] = sum;
  }
  for (int t = threadIdx.x; t < 576; t += blockDim.x) {
    const int i = t / 24, j = t % 24;
    double sn = 0.0, sm = 0.0;
    for (int gp = 0; gp < 8; ++gp) {
      const double *B = bmat_g + gp * 144;
      const double ni = B[i] + B[24 + i] + B[48 + i], nj = B[j] + B[24 + j] + B[48 + j];
      const double mm = 2.0 * (B[i] * B[j] + B[24 + i] * B[24 + j] + B[48 + i] * B[48 + j]) +
                        B[72 + i] * B[72 + j] + B[96 + i] * B[96 + j] + B[120 + i] * B[120 + j];
      sn += g.wg * (ni * nj);
      sm += g.wg * mm;
    }
    kn[t] = sn;
    km[t] = sm;
  }
  __syncthreads();
  const int s = blockIdx.y;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = e < g.nelem;
  int p = 0, node[8];
  double tE[24], tN[24], sE[6], sN[6];
  if (act) {
    const int ex = e % g.nex, ey = (e / g.nex) % g.ney, ez = e / (g.nex * g.ney);
    p = et[e];
    const material_base dE = dmat(mats[p], 0), dN = dmat(mats[p], 1);
    double ue[24];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      node[a] = ((ez + kHexPos[a][2]) * g.ny + (ey + kHexPos[a][1])) * g.nx + ex + kHexPos[a][0];
#pragma unroll
      for (int d = 0; d < 3; ++d) ue[a * 3 + d] = U[(size_t)(d * K + s0 + s) * NP + node[a]];
    }
    double eu[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      double a1 = 0.0;
#pragma unroll
      for (int c = 0; c < 24; ++c) a1 += bw[q * 24 + c] * ue[c];
      eu[q] = a1;
    }
    lin_stress(dE, eu, sE);
    lin_stress(dN, eu, sN);
#pragma unroll
    for (int r = 0; r < 24; ++r) {
      double an = 0.0, am = 0.0;
#pragma unroll
      for (int c = 0; c < 24; ++c) {
        an += kn[r * 24 + c] * ue[c];
        am += km[r * 24 + c] * ue[c];
      }
      tE[r] = dE.lambda * an + dE.mu * am;
      tN[r] = dN.lambda * an + dN.mu * am;
    }
  }
  for (int v = 0; v < 6; ++v) {
    double c6[6] = {0, 0, 0, 0, 0, 0};
    if (act) {
      double dlE = 0.0, dlN = 0.0;
#pragma unroll
      for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const double l = U[(size_t)(d * K + v) * NP + node[a]];
          dlE += l * tE[a * 3 + d];
          dlN += l * tN[a * 3 + d];
        }
      const double cE = sE[v] / g.vol_tot - dlE, cN = sN[v] / g.vol_tot - dlN;
#pragma unroll
      for (int q = 0; q < 3; ++q)
        if (p == q) {
          c6[q * 2] = cE;
          c6[q * 2 + 1] = cN;
        }
    }
    block_sum<6, 256>(c6, lds);
    if (threadIdx.x < 6) part[((size_t)(s * 6 + v) * 6 + threadIdx.x) * gridDim.x + blockIdx.x] = c6[threadIdx.x];
  }
}

// one block per row: out[row] = (sum of part[row * n .. + n)) / div, fixed order
__global__ void __launch_bounds__(256) k_m_rows(const double *part, int n, double div, double *out) {
  __shared__ double lds[4];
  double s[1] = {0.0};
  const double *row = part + (size_t)blockIdx.x * n;
  for (int j = threadIdx.x; j < n; j += 256) s[0] += row[j];
  block_sum<1, 256>(s, lds);
  if (threadIdx.x == 0) out[blockIdx.x] = s[0] / div;
}

// ---------------------------------------------------------------------------
// Mixed-precision solves of an assembled FP64 ELL system (any material law:
// the Newton tangent of a plastic or damaged RVE).  The inner FP32 CG runs on
// the node-block copy of the matrix's interior block; the defect-correction
// residual b - A du is taken from the FP64 ELL matrix itself and du is kept
// in FP64.  Field 0 of a 4-field system carries the solve; fields 1-3 start
// from a zero residual and stop at once.
// ---------------------------------------------------------------------------
__device__ __forceinline__ size_t sell_ix(int row, int slot) {
  return ((size_t)(row >> 6) * kSlots + slot) * 64 + (row & 63);
}

// FP32 node blocks of the interior block of the ELL matrix (ELL slot nb*3+fj
// of row 3n+fi is node-block slot mslot(nb, fi, fj); columns of Dirichlet DOFs and
// the rows of boundary and padding nodes are zero), and the Jacobi weights
// 1 / A_nn(d, d) (FP64 and FP32, zero off the interior)
__global__ void __launch_bounds__(256) k_m_from_ell(Grid g, int NP, const double *vals, float *A, double *kinv64,
                                                    float *kinv) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= NP) return;
  float *base = A + ((size_t)(n >> 6) * kNodeQuads * 64 + (n & 63)) * 4;
  auto put = [&](int s, float v) { base[(size_t)(s >> 2) * 256 + (s & 3)] = v; };
  int i = 0, j = 0, k = 0;
  bool interior = false;
  if (n < g.nn) {
    i = n % g.nx, j = (n / g.nx) % g.ny, k = n / (g.nx * g.ny);
    interior = !node_on_boundary(g, i, j, k);
  }
  for (int nb = 0; nb < 27; ++nb) {
    const int dz = nb / 9 - 1, dy = (nb / 3) % 3 - 1, dx = nb % 3 - 1;
    const bool keep = interior && !node_on_boundary(g, i + dx, j + dy, k + dz);
    for (int fi = 0; fi < 3; ++fi)
      for (int fj = 0; fj < 3; ++fj)
        put(mslot(nb, fi, fj), keep ? (float)vals[sell_ix(3 * n + fi, nb * 3 + fj)] : 0.0f);
  }
  for (int r = 0; r < kRows; ++r) put(r * 28 + 27, 0.0f);  // row pads
  for (int d = 0; d < 3; ++d) {
    const double w = interior ? 1.0 / vals[sell_ix(3 * n + d, 13 * 3 + d)] : 0.0;
    kinv64[(size_t)d * NP + n] = w;
    kinv[(size_t)d * NP + n] = (float)w;
  }
}

// r = b - A du on the interior rows (FP64, the ELL matrix; du is zero on the
// boundary), one thread per row; the FP32 CG start R = r, X = 0, P = kinv R
// in field 0, zero in the other fields; partials as k_m_resid
// (part[q * gridDim.x + block]: q < K the FP64 ||kinv r||^2 of field q,
// K <= q < 2K the (R, P) of field q - K, 2K <= q the FP64 ||r||^2)
template <int K>
__global__ void __launch_bounds__(256) k_m_resid_ell(Grid g, int NP, const int *cols, const double *vals,
                                                     const double *b, const double *du, const double *kinv64,
                                                     const float *kinv, float4 *R, float4 *X, float4 *P,
                                                     double *part) {
  constexpr int KG = K / 4;
  __shared__ double lds[4 * 3 * K];
  const int row = blockIdx.x * 256 + threadIdx.x;
  double zz = 0.0, rz = 0.0, rr = 0.0;
  if (row < g.nrow) {
    const int n = row / 3, d = row % 3;
    const int i = n % g.nx, j = (n / g.nx) % g.ny, k = n / (g.nx * g.ny);
    float r32 = 0.0f, p32 = 0.0f;
    if (!node_on_boundary(g, i, j, k)) {
      const double r64 = b[row] - row_dot_grid(g, vals, du, row);  // = row_dot(cols, ...) bit for bit
      r32 = (float)r64;
      p32 = kinv[(size_t)d * NP + n] * r32;
      const double z64 = kinv64[(size_t)d * NP + n] * r64;
      zz = z64 * z64;
      rz = (double)r32 * (double)p32;
      rr = r64 * r64;
    }
    const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int gq = 0; gq < KG; ++gq) {
      const size_t o = (size_t)(d * KG + gq) * NP + n;
      R[o] = gq ? zero : make_float4(r32, 0.0f, 0.0f, 0.0f);
      X[o] = zero;
      P[o] = gq ? zero : make_float4(p32, 0.0f, 0.0f, 0.0f);
    }
  }
  double v[3 * K];
#pragma unroll
  for (int q = 0; q < 3 * K; ++q) v[q] = 0.0;
  v[0] = zz;
  v[K] = rz;
  v[2 * K] = rr;
  block_sum<3 * K, 256>(v, lds);
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < 3 * K; ++q) part[(size_t)q * gridDim.x + blockIdx.x] = v[q];
}

// du += X (field 0), FP64, one thread per row
template <int K>
__global__ void __launch_bounds__(256) k_m_accum_ell(Grid g, int NP, const float4 *X, double *du) {
  constexpr int KG = K / 4;
  const int row = blockIdx.x * 256 + threadIdx.x;
  if (row >= g.nrow) return;
  const int n = row / 3, d = row % 3;
  du[row] += (double)X[(size_t)(d * KG) * NP + n].x;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static unsigned nblk(size_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

struct MLaunch {
  int NP;         // nodes rounded up to 256
  int per, nb;    // node blocks of 256 per XCD, blocks launched (8 * per)
  int nsl, pers;  // slices of 64 nodes, slices per XCD (8 * pers launched)
  int nbx;        // blocks of 256 float4 per component
};

static MLaunch mshape(const Grid &g) {
  MLaunch s;
  s.NP = (int)round_up((size_t)g.nn, 256);
  s.per = (s.NP / 256 + 7) / 8;
  s.nb = s.per * 8;
  s.nsl = s.NP / 64;
  s.pers = (s.nsl + 7) / 8;
  s.nbx = (int)nblk((size_t)s.NP / 4, 256);
  return s;
}

size_t mixed_matrix_floats(const Grid &g) { return (size_t)mshape(g).NP * kNodeQuads * 4; }
size_t mixed_vector_len(const Grid &g, int K) { return (size_t)mshape(g).NP * 3 * K; }
size_t mixed_partials(const Grid &g, int K) {
  const MLaunch s = mshape(g);
  const size_t resid = std::max((size_t)s.nb, (size_t)g.nrow_pad / 256) * 3 * K;  // k_m_resid / k_m_resid_ell
  const size_t ave = (size_t)nblk(g.nelem, 256) * 6 * K;
  const size_t iter = (size_t)K * s.nsl + 2 * (size_t)3 * K * (s.NP / 256);
  return std::max(std::max(resid, ave), iter);
}
size_t mixed_state_bytes() { return 2 * sizeof(MState) + sizeof(MOuter); }

hipError_t launch_m_asm32(const Grid &g, const int *et, const double *ke, int phase_mode, double *scratch64, float *A,
                          float *kinv, hipStream_t st) {
  const MLaunch s = mshape(g);
  hipLaunchKernelGGL(k_m_kinv, dim3(s.NP / 256), dim3(256), 0, st, g, s.NP, et, ke, phase_mode, scratch64, kinv);
  hipLaunchKernelGGL(k_m_asm32, dim3(s.NP / 256), dim3(256), 0, st, g, s.NP, et, ke, phase_mode, A);
  return hipGetLastError();
}

#define MPP_K_DISPATCH(K, CALL)                     \
  switch (K) {                                      \
    case 4: {                                       \
      constexpr int KK = 4;                         \
      CALL;                                         \
    } break;                                        \
    case 8: {                                       \
      constexpr int KK = 8;                         \
      CALL;                                         \
    } break;                                        \
    case 12: {                                      \
      constexpr int KK = 12;                        \
      CALL;                                         \
    } break;                                        \
    case 16: {                                      \
      constexpr int KK = 16;                        \
      CALL;                                         \
    } break;                                        \
    default:                                        \
      return hipErrorInvalidValue;                  \
  }

hipError_t launch_m_setbc(const Grid &g, int K, const double *eps, double *U, hipStream_t st) {
  const MLaunch s = mshape(g);
  MPP_K_DISPATCH(K, hipLaunchKernelGGL(k_m_setbc<KK>, dim3(nblk(g.nn, 256)), dim3(256), 0, st, g, s.NP, eps, U));
  return hipGetLastError();
}

hipError_t launch_m_load_u(const Grid &g, int K, int f, const double *u, double *U, hipStream_t st) {
  const MLaunch s = mshape(g);
  MPP_K_DISPATCH(K, hipLaunchKernelGGL(k_m_load_u<KK>, dim3(s.NP / 256), dim3(256), 0, st, g, s.NP, f, u, U));
  return hipGetLastError();
}

hipError_t launch_m_store_u(const Grid &g, int K, int f, const double *U, double *u, hipStream_t st) {
  const MLaunch s = mshape(g);
  MPP_K_DISPATCH(K, hipLaunchKernelGGL(k_m_store_u<KK>, dim3(nblk(g.nn, 256)), dim3(256), 0, st, g, s.NP, f, U, u));
  return hipGetLastError();
}

// residual + outer test + inner start state
hipError_t launch_m_resid(const Grid &g, int K, const int *et, const double *ke, const double *U, const double *F,
                          const float *kinv, float *R, float *X, float *P, double *part, int first, double rel,
                          double abs_tol, double tau_min, int max_inner, void *state, hipStream_t st,
                          const int *active) {
  const MLaunch s = mshape(g);
  MState *ms = static_cast<MState *>(state);
  MOuter *mo = reinterpret_cast<MOuter *>(ms + 2);
  MPP_K_DISPATCH(K, {
    hipLaunchKernelGGL(k_m_resid<KK>, dim3(s.nb, KK / 4), dim3(256), 0, st, g, s.NP, s.per, et, ke, U, F, kinv,
                       reinterpret_cast<float4 *>(R), reinterpret_cast<float4 *>(X), reinterpret_cast<float4 *>(P),
                       part);
    hipLaunchKernelGGL(k_m_outer<KK>, dim3(1), dim3(kRedThreads), 0, st, s.nb, part, first, rel, abs_tol, tau_min,
                       max_inner, active, mo, ms);
  });
  return hipGetLastError();
}

// restriction and coarse solve of R into c.ec (state cur: skipped once every field has stopped)
template <int KK>
static void coarse_solve(const Grid &g, const MLaunch &s, const CoarseDev &c, const float *R, const MState *cur,
                         hipStream_t st, bool rx = true) {
  constexpr int KG = KK / 4;
  const CGrid &cg = c.cg;
  const size_t rx_threads = (size_t)3 * KG * (g.ny - 2) * (g.nz - 2) * cg.mx;
  if (rx)  // else T1 is already made (k_m_upd<K, true>)
    hipLaunchKernelGGL(k_c_rx<KK>, dim3(nblk(rx_threads, 256)), dim3(256), 0, st, g, cg, s.NP,
                       reinterpret_cast<const float4 *>(R), reinterpret_cast<float4 *>(c.t1), cur);
  hipLaunchKernelGGL(k_c_ryz<KK>, dim3(cg.my * cg.mz, 3 * KG), dim3(256), 0, st, g, cg,
                     reinterpret_cast<const float4 *>(c.t1), reinterpret_cast<float4 *>(c.rc), cur);
  hipLaunchKernelGGL(k_c_gemv<KK>, dim3(nblk(cg.nc, 64), kCoarseSplit), dim3(256), 0, st, cg, c.ainv, c.rc, c.ecp,
                     cur);
  hipLaunchKernelGGL(k_c_gemv_sum<KK>, dim3(nblk((size_t)cg.nc * KK, 256)), dim3(256), 0, st, cg, c.ecp, c.ec,
                     cur);
}

hipError_t launch_m_iter(const Grid &g, int K, const float *A, const float *kinv, float *R, float *X, float *P,
                         float *AP, double *part, void *state, int parity, hipStream_t st, const CoarseDev *coarse) {
  const MLaunch s = mshape(g);
  MState *ms = static_cast<MState *>(state);
  MState *cur = ms + (parity & 1);
  MState *nxt = ms + ((parity + 1) & 1);
  const int nbx = (int)nblk(s.NP, kVecNodes);
  // with a coarse level k_m_upd runs on blocks of whole x-lines (lpb a block, nbu blocks)
  const int lpb = coarse ? kVecNodes / g.nx : 0;
  const int nbu = coarse ? (g.ny * g.nz + lpb - 1) / lpb : nbx;
  double *part1 = part;                                   // K x nsl
  double *pz = part + (size_t)K * s.nsl;                  // 3K x nbu: (z, z)
  double *pr = pz + (size_t)3 * K * std::max(nbx, nbu);   // 3K x nbx: (r, z)
  const dim3 vgrid(nbx, 3 * K / 4);
  const CGrid cg = coarse ? coarse->cg : CGrid{};
  MPP_K_DISPATCH(K, {
    hipLaunchKernelGGL(k_m_spmv<KK>, dim3(8 * s.pers), dim3(16 * KK), 0, st, g, s.NP, s.nsl, s.pers,
                       reinterpret_cast<const float4 *>(A), reinterpret_cast<const float4 *>(P),
                       reinterpret_cast<float4 *>(AP), part1, cur);
    if (coarse) {
      // R -= alpha AP with the x-restriction of the new R; then the coarse
      // solve, z and (r, z) with the coarse correction (its partials in
      // place of the Jacobi (r, z) ones), and the p-update from z
      hipLaunchKernelGGL((k_m_upd<KK, true>), dim3(nbu, 3 * K / 4), dim3(kRedThreads), 0, st, s.NP, s.nsl, part1,
                         kinv, reinterpret_cast<const float4 *>(AP), reinterpret_cast<float4 *>(R), pz, pr, cur, g,
                         cg, lpb, reinterpret_cast<float4 *>(coarse->t1));
      coarse_solve<KK>(g, s, *coarse, R, cur, st, false);
      hipLaunchKernelGGL((k_c_zdot<KK, false>), vgrid, dim3(kZdotThreads), 0, st, g, cg, s.NP, kinv,
                         reinterpret_cast<const float4 *>(R), reinterpret_cast<const float4 *>(coarse->ec),
                         reinterpret_cast<float4 *>(coarse->z), pr, cur);
      hipLaunchKernelGGL((k_m_pupd<KK, true>), vgrid, dim3(kRedThreads), 0, st, s.NP, kinv,
                         reinterpret_cast<const float4 *>(R), reinterpret_cast<const float4 *>(coarse->z),
                         reinterpret_cast<float4 *>(P), reinterpret_cast<float4 *>(X), pz, nbu, pr, cur, nxt);
    } else {
      hipLaunchKernelGGL((k_m_upd<KK, false>), vgrid, dim3(kRedThreads), 0, st, s.NP, s.nsl, part1, kinv,
                         reinterpret_cast<const float4 *>(AP), reinterpret_cast<float4 *>(R), pz, pr, cur, g, cg,
                         0, nullptr);
      hipLaunchKernelGGL((k_m_pupd<KK, false>), vgrid, dim3(kRedThreads), 0, st, s.NP, kinv,
                         reinterpret_cast<const float4 *>(R), nullptr, reinterpret_cast<float4 *>(P),
                         reinterpret_cast<float4 *>(X), pz, nbx, pr, cur, nxt);
    }
  });
  return hipGetLastError();
}

hipError_t launch_c_start(const Grid &g, int K, const CoarseDev &c, const float *kinv, const float *R, float *P,
                          void *state, hipStream_t st) {
  const MLaunch s = mshape(g);
  MState *s0 = static_cast<MState *>(state);
  const int nbx = (int)nblk(s.NP, kVecNodes);
  const dim3 vgrid(nbx, 3 * K / 4);
  MPP_K_DISPATCH(K, {
    coarse_solve<KK>(g, s, c, R, s0, st);
    hipLaunchKernelGGL((k_c_zdot<KK, true>), vgrid, dim3(kZdotThreads), 0, st, g, c.cg, s.NP, kinv,
                       reinterpret_cast<const float4 *>(R), reinterpret_cast<const float4 *>(c.ec),
                       reinterpret_cast<float4 *>(P), c.prz, s0);
    hipLaunchKernelGGL(k_c_rzfix<KK>, dim3(1), dim3(kRedThreads), 0, st, nbx, c.prz, s0);
  });
  return hipGetLastError();
}

size_t coarse_vec_blocks(const Grid &g) { return nblk(mshape(g).NP, kVecNodes); }

hipError_t launch_m_spmv_only(const Grid &g, int K, const float *A, const float *P, float *AP, double *part,
                              const void *state, hipStream_t st) {
  const MLaunch s = mshape(g);
  const MState *cur = static_cast<const MState *>(state);
  MPP_K_DISPATCH(K, hipLaunchKernelGGL(k_m_spmv<KK>, dim3(8 * s.pers), dim3(16 * KK), 0, st, g, s.NP, s.nsl, s.pers,
                                       reinterpret_cast<const float4 *>(A), reinterpret_cast<const float4 *>(P),
                                       reinterpret_cast<float4 *>(AP), part, cur));
  return hipGetLastError();
}

hipError_t launch_m_from_ell(const Grid &g, const double *vals, float *A, double *kinv64, float *kinv,
                             hipStream_t st) {
  const MLaunch s = mshape(g);
  hipLaunchKernelGGL(k_m_from_ell, dim3(s.NP / 256), dim3(256), 0, st, g, s.NP, vals, A, kinv64, kinv);
  return hipGetLastError();
}

hipError_t launch_m_resid_ell(const Grid &g, int K, const int *cols, const double *vals, const double *b,
                              const double *du, const double *kinv64, const float *kinv, float *R, float *X,
                              float *P, double *part, int first, double rel, double abs_tol, double tau_min,
                              int max_inner, void *state, hipStream_t st) {
  const MLaunch s = mshape(g);
  MState *ms = static_cast<MState *>(state);
  MOuter *mo = reinterpret_cast<MOuter *>(ms + 2);
  const int nbr = g.nrow_pad / 256;
  MPP_K_DISPATCH(K, {
    hipLaunchKernelGGL(k_m_resid_ell<KK>, dim3(nbr), dim3(256), 0, st, g, s.NP, cols, vals, b, du, kinv64, kinv,
                       reinterpret_cast<float4 *>(R), reinterpret_cast<float4 *>(X), reinterpret_cast<float4 *>(P),
                       part);
    hipLaunchKernelGGL(k_m_outer<KK>, dim3(1), dim3(kRedThreads), 0, st, nbr, part, first, rel, abs_tol, tau_min,
                       max_inner, nullptr, mo, ms);
  });
  return hipGetLastError();
}

hipError_t launch_m_accum_ell(const Grid &g, int K, const float *X, double *du, hipStream_t st) {
  const MLaunch s = mshape(g);
  MPP_K_DISPATCH(K, hipLaunchKernelGGL(k_m_accum_ell<KK>, dim3(g.nrow_pad / 256), dim3(256), 0, st, g, s.NP,
                                       reinterpret_cast<const float4 *>(X), du));
  return hipGetLastError();
}

hipError_t launch_m_accum(const Grid &g, int K, const float *X, double *U, hipStream_t st) {
  const MLaunch s = mshape(g);
  MPP_K_DISPATCH(K, hipLaunchKernelGGL(k_m_accum<KK>, dim3(s.NP / 256, 3 * KK / 4), dim3(256), 0, st, s.NP,
                                       reinterpret_cast<const float4 *>(X), U));
  return hipGetLastError();
}

hipError_t launch_m_ave(const Grid &g, int K, const double *bmat, const material_base *mats_d, const int *et,
                        const double *U, double *part, double *out, hipStream_t st) {
  const MLaunch s = mshape(g);
  const unsigned nb = nblk(g.nelem, 256);
  MPP_K_DISPATCH(K, hipLaunchKernelGGL(k_m_ave<KK>, dim3(nb), dim3(256), 0, st, g, s.NP, bmat, mats_d, et, U, part));
  hipLaunchKernelGGL(k_m_rows, dim3(6 * K), dim3(256), 0, st, part, (int)nb, g.vol_tot, out);
  return hipGetLastError();
}

size_t mixed_adj_partials(const Grid &g, int ns) {
  return (size_t)nblk(g.nelem, 256) * 36 * (size_t)std::max(1, ns);
}

hipError_t launch_m_adj_src(const Grid &g, int K, int nadj, const double *bmat, const material_base *mats_d,
                            const int *et, double *G, double *F, hipStream_t st) {
  const MLaunch s = mshape(g);
  hipLaunchKernelGGL(k_m_adj_G, dim3(2), dim3(256), 0, st, bmat, mats_d, g.wg, G);
  MPP_K_DISPATCH(K, hipLaunchKernelGGL(k_m_adj_src<KK>, dim3(s.NP / 256), dim3(256), 0, st, g, s.NP, nadj, et, G,
                                       F));
  return hipGetLastError();
}

hipError_t launch_m_adj(const Grid &g, int K, const double *bmat, const material_base *mats_d, const int *et,
                        const double *ke, const double *U, int s0, int ns, double *part, double *ctan, double *dmat,
                        hipStream_t st) {
  const MLaunch s = mshape(g);
  const unsigned nb = nblk(g.nelem, 256), nbb = nblk(boundary_elems(g), 256);
  MPP_K_DISPATCH(K, hipLaunchKernelGGL(k_m_adj_ctan<KK>, dim3(nbb, 6), dim3(256), 0, st, g, s.NP, bmat, mats_d, et,
                                       ke, U, part));
  hipLaunchKernelGGL(k_m_rows, dim3(36), dim3(256), 0, st, part, (int)nbb, 1.0, ctan);
  if (ns > 0) {
    MPP_K_DISPATCH(K, hipLaunchKernelGGL(k_m_adj_dmat<KK>, dim3(nb, ns), dim3(256), 0, st, g, s.NP, bmat,
                                         mats_d, et, U, s0, part));
    hipLaunchKernelGGL(k_m_rows, dim3(36 * ns), dim3(256), 0, st, part, (int)nb, 1.0, dmat);
  }
  return hipGetLastError();
}

// host views of the state (offsets into the state block)
void mixed_state_view(const void *state_h, int parity, int *its, int *all_done) {
  const MState *ms = static_cast<const MState *>(state_h) + (parity & 1);
  *its = ms->its;
  *all_done = ms->all_done;
}
int mixed_state_field(const void *state_h, int parity, int k, double *pn, double *target) {
  const MState *ms = static_cast<const MState *>(state_h) + (parity & 1);
  *pn = ms->pn[k];
  *target = ms->tau[k] * ms->pn0[k];
  return ms->done[k];
}
void mixed_outer_view(const void *state_h, int K, int *all_conv, double *pn, double *pn0) {
  const MOuter *mo = reinterpret_cast<const MOuter *>(static_cast<const MState *>(state_h) + 2);
  *all_conv = mo->all_conv;
  for (int r = 0; r < K; ++r) {
    pn[r] = mo->pn[r];
    pn0[r] = mo->pn0[r];
  }
}
void mixed_outer_norms(const void *state_h, int K, int *conv, double *rn, double *rn0) {
  const MOuter *mo = reinterpret_cast<const MOuter *>(static_cast<const MState *>(state_h) + 2);
  for (int r = 0; r < K; ++r) {
    conv[r] = mo->conv[r];
    rn[r] = mo->rn[r];
    rn0[r] = mo->rn0[r];
  }
}

}  // namespace mpp
