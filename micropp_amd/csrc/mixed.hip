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
// FP32 matrix layout (node-blocked SELL-64): one thread per node; the node's
// 27 neighbour blocks of 3x3 coefficients (slot s = nb*9 + fi*3 + fj, nb as
// in the reference ELL) are padded to 244 floats and stored as 61 float4 per
// node, quad-major inside a slice of 64 nodes: quad q of node n sits at
// float4 index ((n / 64) * 61 + q) * 64 + n % 64.  Column indices are implicit
// (structured 27-point stencil).  Rows and columns of Dirichlet nodes are
// dropped (the CG vectors are zero there), which leaves the symmetric interior
// block.  Vectors are [node][3][K] so a neighbour gather reads 3K consecutive
// floats.
//
// Reductions are fence-free: blocks write FP64 partials and the next kernel
// re-sums them in a fixed order.  Iteration k of the inner CG reads
// state[k & 1] and writes state[(k+1) & 1].  Row-space kernels map blocks to
// XCDs so that each XCD's L2 holds a contiguous slab of nodes.
#include "internal.hpp"
#include "physics.hpp"
#include "reduce.hpp"

namespace mpp {

constexpr int kMaxRhs = 16;
constexpr int kNodeQuads = 61;  // 244 floats per node (243 coefficients + 1 pad)

struct MState {
  double rz[kMaxRhs], pn[kMaxRhs], pn0[kMaxRhs], tau[kMaxRhs];
  int done[kMaxRhs];
  int its, max_its, all_done, pad;
};

struct MOuter {
  double pn[kMaxRhs], pn0[kMaxRhs], target[kMaxRhs];
  int conv[kMaxRhs];
  int all_conv, pad[3];
};

__device__ __forceinline__ bool node_on_boundary(const Grid &g, int i, int j, int k) {
  return i == 0 || j == 0 || k == 0 || i == g.nx - 1 || j == g.ny - 1 || k == g.nz - 1;
}

// Blocks are dispatched round-robin over the 8 XCDs; launch 8 * per blocks and
// give XCD x the logical blocks [x * per, (x + 1) * per).
__device__ __forceinline__ int xcd_block(int per) {
  const int b = blockIdx.x;
  return (b & 7) * per + (b >> 3);
}

// ---------------------------------------------------------------------------
// FP32 matrix from the element blocks (phase table or per element), and the
// FP32 Jacobi weights.  Per slot, the element contributions are summed in
// FP64 in the reference element order, then rounded once.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_m_asm32(Grid g, int nn_pad, const int *et, const double *ke, int phase_mode,
                                                 float *A, float *kinv) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn_pad) return;
  float *base = A + ((size_t)(n >> 6) * kNodeQuads * 64 + (n & 63)) * 4;
  auto put = [&](int s, float v) { base[(size_t)(s >> 2) * 256 + (s & 3)] = v; };
  for (int s = 0; s < kNodeQuads * 4; ++s) put(s, 0.0f);
  if (n >= g.nn) {
    for (int d = 0; d < 3; ++d) kinv[(size_t)n * 3 + d] = 0.0f;
    return;
  }
  const int i = n % g.nx, j = (n / g.nx) % g.ny, k = n / (g.nx * g.ny);
  if (node_on_boundary(g, i, j, k)) {
    for (int d = 0; d < 3; ++d) {
      put(13 * 9 + d * 3 + d, 1.0f);
      kinv[(size_t)n * 3 + d] = 0.0f;  // keeps z, p and Ap zero on Dirichlet DOFs
    }
    return;
  }
  const double *blk[8];
  int la[8];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int c = 0; c < 2; ++c) {
        const int q = (a * 2 + b) * 2 + c;
        const int e = ((k - 1 + c) * g.ney + (j - 1 + b)) * g.nex + (i - 1 + a);
        blk[q] = ke + (size_t)(phase_mode ? et[e] : e) * 576;
        la[q] = local_node(1 - a, 1 - b, 1 - c);
      }
  for (int nb = 0; nb < 27; ++nb) {
    const int dz = nb / 9 - 1, dy = (nb / 3) % 3 - 1, dx = nb % 3 - 1;
    if (node_on_boundary(g, i + dx, j + dy, k + dz)) continue;  // column of a Dirichlet DOF: dropped
    for (int fi = 0; fi < 3; ++fi)
      for (int fj = 0; fj < 3; ++fj) {
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
        put(nb * 9 + fi * 3 + fj, (float)acc);
        if (nb == 13 && fi == fj) kinv[(size_t)n * 3 + fi] = (float)(1.0 / acc);
      }
  }
}

// ---------------------------------------------------------------------------
// boundary values of K fields (set_displ_bc, src/micro3D.cpp:26-78), FP64
// ---------------------------------------------------------------------------
template <int K>
__global__ void __launch_bounds__(256) k_m_setbc(Grid g, const double *eps, double *U) {
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
      U[((size_t)n * 3 + d) * K + r] = s;
    }
  }
}

// ---------------------------------------------------------------------------
// FP64 residual r = -(K u) on interior DOFs (elastic: element blocks of the
// phase table), rounded into the FP32 CG start: R = r, X = 0, P = kinv R.
// Partials per block: [ (kinv r)^2 summed in FP64 | R.P ] for each field.
// ---------------------------------------------------------------------------
template <int K>
__global__ void __launch_bounds__(256) k_m_resid(Grid g, int nn_pad, int per, const int *et, const double *ke,
                                                 const double *U, const float *kinv, float *R, float *X, float *P,
                                                 double *part) {
  __shared__ double ke_s[3 * 576];
  __shared__ double lds[4 * 2 * K];
  for (int t = threadIdx.x; t < 3 * 576; t += blockDim.x) ke_s[t] = ke[t];
  __syncthreads();
  const int lb = xcd_block(per);
  const int n = lb * 256 + threadIdx.x;
  double zz[K], rz[K];
#pragma unroll
  for (int r = 0; r < K; ++r) zz[r] = rz[r] = 0.0;
  if (n < nn_pad) {
    const int i = n % g.nx, j = (n / g.nx) % g.ny, k = n / (g.nx * g.ny);
    const bool interior = n < g.nn && !node_on_boundary(g, i, j, k);
#pragma unroll
    for (int kg = 0; kg < K; kg += 4) {
      double acc[3][4];
#pragma unroll
      for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[d][q] = 0.0;
      if (interior) {
        for (int a = 0; a < 2; ++a)
          for (int b = 0; b < 2; ++b)
            for (int c = 0; c < 2; ++c) {
              const int ex = i - 1 + a, ey = j - 1 + b, ez = k - 1 + c;
              const int e = (ez * g.ney + ey) * g.nex + ex;
              const double *kb = ke_s + et[e] * 576 + local_node(1 - a, 1 - b, 1 - c) * 3 * 24;
              for (int m = 0; m < 8; ++m) {
                const int node = ((ez + kHexPos[m][2]) * g.ny + (ey + kHexPos[m][1])) * g.nx + ex + kHexPos[m][0];
                const double *u = U + (size_t)node * 3 * K + kg;
#pragma unroll
                for (int fj = 0; fj < 3; ++fj) {
                  const double2 u01 = *reinterpret_cast<const double2 *>(u + fj * K);
                  const double2 u23 = *reinterpret_cast<const double2 *>(u + fj * K + 2);
#pragma unroll
                  for (int d = 0; d < 3; ++d) {
                    const double cf = kb[d * 24 + m * 3 + fj];
                    acc[d][0] += cf * u01.x;
                    acc[d][1] += cf * u01.y;
                    acc[d][2] += cf * u23.x;
                    acc[d][3] += cf * u23.y;
                  }
                }
              }
            }
      }
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const float kv = kinv[(size_t)n * 3 + d];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const double r64 = -acc[d][q];
          const float r32 = (float)r64;
          const float p32 = kv * r32;
          const size_t idx = ((size_t)n * 3 + d) * K + kg + q;
          R[idx] = r32;
          X[idx] = 0.0f;
          P[idx] = p32;
          const double z64 = (double)kv * r64;
          zz[kg + q] += z64 * z64;
          rz[kg + q] += (double)r32 * (double)p32;
        }
      }
    }
  }
  double v[2 * K];
#pragma unroll
  for (int r = 0; r < K; ++r) {
    v[r] = zz[r];
    v[K + r] = rz[r];
  }
  block_sum<2 * K, 256>(v, lds);
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < 2 * K; ++q) part[(size_t)blockIdx.x * 2 * K + q] = v[q];
}

template <int NV>
__device__ __forceinline__ void sum_parts(const double *part, int nparts, int stride, double (&out)[NV], double *lds) {
  double acc[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) acc[q] = 0.0;
  for (int b = threadIdx.x; b < nparts; b += 256)
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] += part[(size_t)b * stride + q];
  block_sum<NV, 256>(acc, lds);
#pragma unroll
  for (int q = 0; q < NV; ++q) out[q] = acc[q];
}

// one block: outer convergence test and the inner CG start state
template <int K>
__global__ void __launch_bounds__(256) k_m_outer(int nparts, const double *part, int first, double rel, double abs_tol,
                                                 double tau_min, int max_inner, MOuter *o, MState *s0) {
  __shared__ double lds[4 * 2 * K];
  double t[2 * K];
  sum_parts<2 * K>(part, nparts, 2 * K, t, lds);
  if (threadIdx.x != 0) return;
  int all = 1;
#pragma unroll
  for (int r = 0; r < K; ++r) {
    const double pn = sqrt(t[r]);
    if (first) {
      o->pn0[r] = pn;
      o->target[r] = fmax(abs_tol, rel * pn);
    }
    o->pn[r] = pn;
    const int conv = pn < o->target[r] || pn == 0.0;
    o->conv[r] = conv;
    all &= conv;
    s0->rz[r] = t[K + r];
    s0->pn[r] = pn;
    s0->pn0[r] = pn;
    // aim a little below what the outer test needs, but not below what FP32 can resolve
    s0->tau[r] = (pn > 0.0) ? fmax(tau_min, 0.3 * o->target[r] / pn) : 1.0;
    s0->done[r] = conv;
  }
  o->all_conv = all;
  s0->its = 0;
  s0->max_its = max_inner;
  s0->all_done = all;
}

// ---------------------------------------------------------------------------
// inner CG, iteration part 1: AP = A P and partials of P.AP (FP64)
// ---------------------------------------------------------------------------
template <int K>
__global__ void __launch_bounds__(256) k_m_spmv(Grid g, int nn_pad, int per, const float4 *A, const float *P, float *AP,
                                                double *part, const MState *cur) {
  __shared__ double lds[4 * K];
  if (cur->all_done) return;
  const int lb = xcd_block(per);
  const int n = lb * 256 + threadIdx.x;
  double pap[K];
#pragma unroll
  for (int r = 0; r < K; ++r) pap[r] = 0.0;
  if (n < nn_pad) {
    float acc[3][K];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int r = 0; r < K; ++r) acc[d][r] = 0.0f;
    if (n < g.nn) {
      const int i = n % g.nx, j = (n / g.nx) % g.ny, k = n / (g.nx * g.ny);
      const int nxny = g.nx * g.ny;
      typedef float f4v __attribute__((ext_vector_type(4)));
      const f4v *arow = reinterpret_cast<const f4v *>(A) + (size_t)(n >> 6) * kNodeQuads * 64 + (n & 63);
      // 7 chunks of up to 4 neighbours = 9 quads each (the last: 3 neighbours,
      // 7 quads); the chunk loop stays rolled to bound the live registers
#pragma unroll 1
      for (int ch = 0; ch < 7; ++ch) {
        constexpr int kq = 9;
        float a[36];
#pragma unroll
        for (int q = 0; q < kq; ++q) {
          if (ch * 9 + q < kNodeQuads) {
            const f4v v = __builtin_nontemporal_load(&arow[(size_t)(ch * 9 + q) * 64]);
            a[q * 4 + 0] = v.x;
            a[q * 4 + 1] = v.y;
            a[q * 4 + 2] = v.z;
            a[q * 4 + 3] = v.w;
          }
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int nb = ch * 4 + t;
          if (nb < 27) {
            const int dz = nb / 9 - 1, dy = (nb / 3) % 3 - 1, dx = nb % 3 - 1;
            const bool in = (i + dx >= 0 && i + dx < g.nx) && (j + dy >= 0 && j + dy < g.ny) &&
                            (k + dz >= 0 && k + dz < g.nz);
            const int m = in ? n + dz * nxny + dy * g.nx + dx : n;  // outside: zero block, any valid node
            const float *pm = P + (size_t)m * 3 * K;
#pragma unroll
            for (int fj = 0; fj < 3; ++fj) {
              float pv[K];
#pragma unroll
              for (int r = 0; r < K; r += 4) {
                const float4 v = *reinterpret_cast<const float4 *>(pm + fj * K + r);
                pv[r] = v.x;
                pv[r + 1] = v.y;
                pv[r + 2] = v.z;
                pv[r + 3] = v.w;
              }
#pragma unroll
              for (int fi = 0; fi < 3; ++fi) {
                const float c = a[t * 9 + fi * 3 + fj];
#pragma unroll
                for (int r = 0; r < K; ++r) acc[fi][r] += c * pv[r];
              }
            }
          }
        }
      }
    }
    float *ap = AP + (size_t)n * 3 * K;
    const float *pn = P + (size_t)n * 3 * K;
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int r = 0; r < K; r += 4) {
        *reinterpret_cast<float4 *>(ap + d * K + r) = make_float4(acc[d][r], acc[d][r + 1], acc[d][r + 2], acc[d][r + 3]);
        const float4 pv = *reinterpret_cast<const float4 *>(pn + d * K + r);
        pap[r] += (double)pv.x * acc[d][r];
        pap[r + 1] += (double)pv.y * acc[d][r + 1];
        pap[r + 2] += (double)pv.z * acc[d][r + 2];
        pap[r + 3] += (double)pv.w * acc[d][r + 3];
      }
  }
  block_sum<K, 256>(pap, lds);
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < K; ++q) part[(size_t)blockIdx.x * K + q] = pap[q];
}

// part 2: alpha; X += alpha P; R -= alpha AP; partials of z.z and R.z (z = kinv R)
template <int K>
__global__ void __launch_bounds__(256) k_m_upd(int nn_pad, int per, int nparts, const double *part1, const float *P,
                                               const float *AP, const float *kinv, float *X, float *R, double *part2,
                                               const MState *cur) {
  __shared__ double lds[4 * 2 * K];
  __shared__ float alpha_s[K];
  if (cur->all_done) return;
  double pap[K];
  sum_parts<K>(part1, nparts, K, pap, lds);
  if (threadIdx.x == 0)
#pragma unroll
    for (int r = 0; r < K; ++r) alpha_s[r] = cur->done[r] ? 0.0f : (float)(cur->rz[r] / pap[r]);
  __syncthreads();
  const int lb = xcd_block(per);
  const int n = lb * 256 + threadIdx.x;
  double zz[K], rz[K];
#pragma unroll
  for (int r = 0; r < K; ++r) zz[r] = rz[r] = 0.0;
  if (n < nn_pad) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const float kv = kinv[(size_t)n * 3 + d];
      const size_t o = ((size_t)n * 3 + d) * K;
#pragma unroll
      for (int r = 0; r < K; r += 4) {
        const float4 p = *reinterpret_cast<const float4 *>(P + o + r);
        const float4 ap = *reinterpret_cast<const float4 *>(AP + o + r);
        float4 x = *reinterpret_cast<const float4 *>(X + o + r);
        float4 rr = *reinterpret_cast<const float4 *>(R + o + r);
        x.x += alpha_s[r] * p.x;
        x.y += alpha_s[r + 1] * p.y;
        x.z += alpha_s[r + 2] * p.z;
        x.w += alpha_s[r + 3] * p.w;
        rr.x -= alpha_s[r] * ap.x;
        rr.y -= alpha_s[r + 1] * ap.y;
        rr.z -= alpha_s[r + 2] * ap.z;
        rr.w -= alpha_s[r + 3] * ap.w;
        *reinterpret_cast<float4 *>(X + o + r) = x;
        *reinterpret_cast<float4 *>(R + o + r) = rr;
        const float z0 = kv * rr.x, z1 = kv * rr.y, z2 = kv * rr.z, z3 = kv * rr.w;
        zz[r] += (double)z0 * z0;
        zz[r + 1] += (double)z1 * z1;
        zz[r + 2] += (double)z2 * z2;
        zz[r + 3] += (double)z3 * z3;
        rz[r] += (double)rr.x * z0;
        rz[r + 1] += (double)rr.y * z1;
        rz[r + 2] += (double)rr.z * z2;
        rz[r + 3] += (double)rr.w * z3;
      }
    }
  }
  double v[2 * K];
#pragma unroll
  for (int r = 0; r < K; ++r) {
    v[r] = zz[r];
    v[K + r] = rz[r];
  }
  block_sum<2 * K, 256>(v, lds);
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < 2 * K; ++q) part2[(size_t)blockIdx.x * 2 * K + q] = v[q];
}

// part 3: beta; P = kinv R + beta P; block 0 writes the next state
template <int K>
__global__ void __launch_bounds__(256) k_m_pupd(int nn_pad, int per, int nparts, const double *part2, const float *R,
                                                const float *kinv, float *P, const MState *cur, MState *nxt) {
  __shared__ double lds[4 * 2 * K];
  __shared__ float beta_s[K];
  if (cur->all_done) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *nxt = *cur;
    return;
  }
  double t[2 * K];
  sum_parts<2 * K>(part2, nparts, 2 * K, t, lds);
  if (threadIdx.x == 0)
#pragma unroll
    for (int r = 0; r < K; ++r) beta_s[r] = cur->done[r] ? 0.0f : (float)(t[K + r] / cur->rz[r]);
  __syncthreads();
  const int lb = xcd_block(per);
  const int n = lb * 256 + threadIdx.x;
  if (n < nn_pad) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const float kv = kinv[(size_t)n * 3 + d];
      const size_t o = ((size_t)n * 3 + d) * K;
#pragma unroll
      for (int r = 0; r < K; r += 4) {
        const float4 rr = *reinterpret_cast<const float4 *>(R + o + r);
        float4 p = *reinterpret_cast<const float4 *>(P + o + r);
        p.x = kv * rr.x + beta_s[r] * p.x;
        p.y = kv * rr.y + beta_s[r + 1] * p.y;
        p.z = kv * rr.z + beta_s[r + 2] * p.z;
        p.w = kv * rr.w + beta_s[r + 3] * p.w;
        *reinterpret_cast<float4 *>(P + o + r) = p;
      }
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    int all = 1;
#pragma unroll
    for (int r = 0; r < K; ++r) {
      const int was = cur->done[r];
      const double pn = was ? cur->pn[r] : sqrt(t[r]);
      const int dn = was || pn < cur->tau[r] * cur->pn0[r] || !(pn == pn);
      nxt->rz[r] = was ? cur->rz[r] : t[K + r];
      nxt->pn[r] = pn;
      nxt->pn0[r] = cur->pn0[r];
      nxt->tau[r] = cur->tau[r];
      nxt->done[r] = dn;
      all &= dn;
    }
    const int its = cur->its + 1;
    nxt->its = its;
    nxt->max_its = cur->max_its;
    nxt->all_done = all || its >= cur->max_its;
  }
}

// U += X (FP64 update of the Newton / defect-correction step)
__global__ void __launch_bounds__(256) k_m_accum(size_t n, const float *X, double *U) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) U[t] += (double)X[t];
}

// ---------------------------------------------------------------------------
// calc_ave_stress (src/average.cpp:57-82) of the K fields, FP64: per block
// partials [K][6], then a one-block sum.
// ---------------------------------------------------------------------------
template <int K>
__global__ void __launch_bounds__(256) k_m_ave(Grid g, const double *bmat_g, const material_base *mats, const int *et,
                                               const double *U, double *part) {
  __shared__ double bm[8 * 144];
  __shared__ double lds[4 * 6];
  for (int t = threadIdx.x; t < 8 * 144; t += blockDim.x) bm[t] = bmat_g[t];
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
      for (int m = 0; m < 8; ++m) {
        const int node = ((ez + kHexPos[m][2]) * g.ny + (ey + kHexPos[m][1])) * g.nx + ex + kHexPos[m][0];
        for (int d = 0; d < 3; ++d) ue[m * 3 + d] = U[((size_t)node * 3 + d) * K + r];
      }
      const material_base m = mats[ph];
      for (int gp = 0; gp < 8; ++gp) {
        double eps[6], sig[6];
        for (int v = 0; v < 6; ++v) {
          double s = 0;
          for (int q = 0; q < 24; ++q) s += bm[gp * 144 + v * 24 + q] * ue[q];
          eps[v] = s;
        }
        mat_stress(m, eps, nullptr, sig);
        for (int v = 0; v < 6; ++v) s6[v] += sig[v] * g.wg;
      }
    }
    block_sum<6, 256>(s6, lds);
    if (threadIdx.x < 6) part[((size_t)blockIdx.x * K + r) * 6 + threadIdx.x] = s6[threadIdx.x];
  }
}

__global__ void __launch_bounds__(256) k_m_ave_finish(int nparts, int nv, const double *part, double vol,
                                                      double *out) {
  // nv = 6 K <= 96: one thread per output, summing its partials in order
  const int q = threadIdx.x;
  if (q >= nv) return;
  double s = 0.0;
  for (int b = 0; b < nparts; ++b) s += part[(size_t)b * nv + q];
  out[q] = s / vol;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static unsigned nblk(size_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

struct MLaunch {
  int nn_pad, per, nb_launch;
};

static MLaunch mshape(const Grid &g) {
  MLaunch s;
  s.nn_pad = (int)round_up((size_t)g.nn, 256);
  const int nb = s.nn_pad / 256;
  s.per = (nb + 7) / 8;
  s.nb_launch = s.per * 8;
  return s;
}

size_t mixed_matrix_floats(const Grid &g) { return (size_t)mshape(g).nn_pad * kNodeQuads * 4; }
size_t mixed_vector_len(const Grid &g, int K) { return (size_t)mshape(g).nn_pad * 3 * K; }
size_t mixed_partials(const Grid &g, int K) {
  const MLaunch s = mshape(g);
  const size_t a = (size_t)s.nb_launch * 2 * K;
  const size_t b = (size_t)nblk(g.nelem, 256) * 6 * K;
  return std::max(a, b) + (size_t)s.nb_launch * K;
}
size_t mixed_state_bytes() { return 2 * sizeof(MState) + sizeof(MOuter); }

hipError_t launch_m_asm32(const Grid &g, const int *et, const double *ke, int phase_mode, float *A, float *kinv,
                          hipStream_t st) {
  const MLaunch s = mshape(g);
  hipLaunchKernelGGL(k_m_asm32, dim3(s.nn_pad / 256), dim3(256), 0, st, g, s.nn_pad, et, ke, phase_mode, A, kinv);
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
  MPP_K_DISPATCH(K, hipLaunchKernelGGL(k_m_setbc<KK>, dim3(nblk(g.nn, 256)), dim3(256), 0, st, g, eps, U));
  return hipGetLastError();
}

// residual + outer test + inner start state
hipError_t launch_m_resid(const Grid &g, int K, const int *et, const double *ke, const double *U, const float *kinv,
                          float *R, float *X, float *P, double *part, int first, double rel, double abs_tol,
                          double tau_min, int max_inner, void *state, hipStream_t st) {
  const MLaunch s = mshape(g);
  MState *ms = static_cast<MState *>(state);
  MOuter *mo = reinterpret_cast<MOuter *>(ms + 2);
  MPP_K_DISPATCH(K, {
    hipLaunchKernelGGL(k_m_resid<KK>, dim3(s.nb_launch), dim3(256), 0, st, g, s.nn_pad, s.per, et, ke, U, kinv, R, X,
                       P, part);
    hipLaunchKernelGGL(k_m_outer<KK>, dim3(1), dim3(256), 0, st, s.nb_launch, part, first, rel, abs_tol, tau_min,
                       max_inner, mo, ms);
  });
  return hipGetLastError();
}

hipError_t launch_m_iter(const Grid &g, int K, const float *A, float *X, float *R, float *P, float *AP,
                         const float *kinv, double *part, void *state, int parity, hipStream_t st) {
  const MLaunch s = mshape(g);
  MState *ms = static_cast<MState *>(state);
  const MState *cur = ms + (parity & 1);
  MState *nxt = ms + ((parity + 1) & 1);
  double *part1 = part;                                // nb_launch * K
  double *part2 = part + (size_t)s.nb_launch * K;      // nb_launch * 2K
  MPP_K_DISPATCH(K, {
    hipLaunchKernelGGL(k_m_spmv<KK>, dim3(s.nb_launch), dim3(256), 0, st, g, s.nn_pad, s.per,
                       reinterpret_cast<const float4 *>(A), P, AP, part1, cur);
    hipLaunchKernelGGL(k_m_upd<KK>, dim3(s.nb_launch), dim3(256), 0, st, s.nn_pad, s.per, s.nb_launch, part1, P, AP,
                       kinv, X, R, part2, cur);
    hipLaunchKernelGGL(k_m_pupd<KK>, dim3(s.nb_launch), dim3(256), 0, st, s.nn_pad, s.per, s.nb_launch, part2, R,
                       kinv, P, cur, nxt);
  });
  return hipGetLastError();
}

hipError_t launch_m_spmv_only(const Grid &g, int K, const float *A, const float *P, float *AP, double *part,
                              const void *state, hipStream_t st) {
  const MLaunch s = mshape(g);
  const MState *cur = static_cast<const MState *>(state);
  MPP_K_DISPATCH(K, hipLaunchKernelGGL(k_m_spmv<KK>, dim3(s.nb_launch), dim3(256), 0, st, g, s.nn_pad, s.per,
                                       reinterpret_cast<const float4 *>(A), P, AP, part, cur));
  return hipGetLastError();
}

hipError_t launch_m_accum(const Grid &g, int K, const float *X, double *U, hipStream_t st) {
  const size_t n = mixed_vector_len(g, K);
  hipLaunchKernelGGL(k_m_accum, dim3(nblk(n, 256)), dim3(256), 0, st, n, X, U);
  return hipGetLastError();
}

hipError_t launch_m_ave(const Grid &g, int K, const double *bmat, const material_base *mats_d, const int *et,
                        const double *U, double *part, double *out, hipStream_t st) {
  const unsigned nb = nblk(g.nelem, 256);
  MPP_K_DISPATCH(K, hipLaunchKernelGGL(k_m_ave<KK>, dim3(nb), dim3(256), 0, st, g, bmat, mats_d, et, U, part));
  hipLaunchKernelGGL(k_m_ave_finish, dim3(1), dim3(128), 0, st, (int)nb, 6 * K, part, g.vol_tot, out);
  return hipGetLastError();
}

// host views of the state (offsets into the state block)
void mixed_state_view(const void *state_h, int parity, int *its, int *all_done) {
  const MState *ms = static_cast<const MState *>(state_h) + (parity & 1);
  *its = ms->its;
  *all_done = ms->all_done;
}
void mixed_outer_view(const void *state_h, int K, int *all_conv, double *pn, double *pn0) {
  const MOuter *mo = reinterpret_cast<const MOuter *>(static_cast<const MState *>(state_h) + 2);
  *all_conv = mo->all_conv;
  for (int r = 0; r < K; ++r) {
    pn[r] = mo->pn[r];
    pn0[r] = mo->pn0[r];
  }
}

}  // namespace mpp
