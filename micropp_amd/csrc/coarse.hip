// coarse.hip -- set-up of the two-level preconditioner of the multi-RHS inner
// CG (the kernels that apply it are mixed.hip k_c_*).
//
// The coarse space: trilinear hats of the interior nodes of a coarse grid
// aligned with the RVE's grid (h fine elements per coarse element and side),
// the Dirichlet boundary excluded.  A_c = P^T A P (Galerkin) is assembled in
// FP64 on the device, element by element: a fine element at local position
// (px, py, pz) inside its coarse element interpolates its 8 nodes from that
// coarse element's 8 corners, so its contribution is C = P_e^T Ke P_e with P_e
// the 24 x 24 map of that position, the same for every element of one phase
// and position (k_cs_table); the coarse element matrices sum those
// (k_cs_elem) and the band tiles of A_c gather them (k_cs_band).  A_c is
// factored and inverted on the device by a banded tile Cholesky that forms
// W = L^-1 in the same launches (k_cs_chol), then A_c^-1 = W^T W (k_cs_wtw),
// kept as a dense FP32 inverse, so one application of the coarse solve is one
// matrix-vector product per field.
#include <algorithm>
#include <cstdlib>
#include <utility>

#include "internal.hpp"

namespace mpp {

namespace {

// coarse elements along a side of ne fine elements: the largest divisor
// d <= dmax of ne with ne / d >= 2 (0: none with d >= 2)
int pick_side(int ne, int dmax) {
  for (int d = std::min(dmax, ne / 2); d >= 2; --d)
    if (ne % d == 0) return d;
  return 0;
}

// corner A of a hexahedron (the element's local node order: (0,0,0), (1,0,0),
// (1,1,0), (0,1,0), then the same at z + 1) and its inverse
__device__ __forceinline__ int hex_ox(int A) { return ((A + 1) >> 1) & 1; }
__device__ __forceinline__ int hex_oy(int A) { return (A >> 1) & 1; }
__device__ __forceinline__ int hex_oz(int A) { return A >> 2; }
__device__ __forceinline__ int hex_corner(int ox, int oy, int oz) { return oz * 4 + (oy ? 3 - ox : ox); }

// C[(pos * 3 + phase) * 576 + r * 24 + c] = (P_e^T Ke_phase P_e)[r][c] for the
// fine element at position pos = (pz hy + py) hx + px: block (pos, phase),
// one thread per entry
__global__ void __launch_bounds__(576) k_cs_table(CGrid cg, const double *ke, double *C) {
  __shared__ double N[8][8];  // N[a][A]: coarse corner A's hat at the fine element's node a
  __shared__ double K[576];
  const int pos = blockIdx.x, ph = blockIdx.y, t = threadIdx.x;
  const int px = pos % cg.hx, py = (pos / cg.hx) % cg.hy, pz = pos / (cg.hx * cg.hy);
  K[t] = ke[(size_t)ph * 576 + t];
  if (t < 64) {
    const int a = t >> 3, A = t & 7;
    const double xi = (double)(px + hex_ox(a)) / cg.hx, yi = (double)(py + hex_oy(a)) / cg.hy,
                 zi = (double)(pz + hex_oz(a)) / cg.hz;
    N[a][A] = (hex_ox(A) ? xi : 1.0 - xi) * (hex_oy(A) ? yi : 1.0 - yi) * (hex_oz(A) ? zi : 1.0 - zi);
  }
  __syncthreads();
  const int A = t / 72, d = (t / 24) % 3, B = (t % 24) / 3, e = t % 3;
  double s = 0.0;
  for (int a = 0; a < 8; ++a) {
    double ta = 0.0;
    for (int b = 0; b < 8; ++b) ta += K[(a * 3 + d) * 24 + b * 3 + e] * N[b][B];
    s += N[a][A] * ta;
  }
  C[((size_t)pos * 3 + ph) * 576 + t] = s;
}

// KE[E * 576 + t]: the stiffness of coarse element E, the sum of its fine
// elements' C entries in position order; block E, one thread per entry
__global__ void __launch_bounds__(576) k_cs_elem(Grid g, CGrid cg, int dx, int dy, const int *et, const double *C,
                                                 double *KE) {
  const int E = blockIdx.x, t = threadIdx.x;
  const int Ex = E % dx, Ey = (E / dx) % dy, Ez = E / (dx * dy);
  double s = 0.0;
  for (int pz = 0; pz < cg.hz; ++pz)
    for (int py = 0; py < cg.hy; ++py)
      for (int px = 0; px < cg.hx; ++px) {
        const int e = ((Ez * cg.hz + pz) * g.ney + (Ey * cg.hy + py)) * g.nex + Ex * cg.hx + px;
        const int pos = (pz * cg.hy + py) * cg.hx + px;
        s += C[((size_t)pos * 3 + et[e]) * 576 + t];
      }
  KE[(size_t)E * 576 + t] = s;
}

// ---------------------------------------------------------------------------
// Factorisation and inverse.  A_c is banded: a coarse node couples only with
// its 26 neighbours, so with the node order above entry (r, c) is zero for
// |r - c| > bw = 3 (mx my + mx + 1) + 2, and so is the Cholesky factor L's.
// Everything works on T x T tiles (kChT = 32, row-major inside a tile, the
// order padded to nt T with the identity): tile (i, j) of A or L is zero for
// i - j > bt = floor((bw + T - 1) / T).  Band tiles are stored by block
// column: slot j (bt + 1) + (i - j).  W = L^-1 is lower triangular and dense:
// tile (i, j), j <= i, in slot i (i + 1) / 2 + j.  Then A_c^-1 = W^T W.
//
// One launch per tile step k (k_cs_chol), k = -1 .. nt - 1:
//  - k = -1: factor tile (0, 0);
//  - the right-looking step k, one block per tile (i, j), k < j <= i <= k +
//    bt: L_ik = A_ik L_kk^-T and L_jk likewise (triangular solves, a lane per
//    row), A_ij -= L_ik L_jk^T; block (i, k + 1) stores L_ik; block (k + 1,
//    k + 1) then factors its updated tile: L_{k+1,k+1};
//  - row k of W, one block per tile (k, j), j <= k: W_kj = L_kk^-1 (delta_kj
//    I - sum_{m = max(j, k - bt)}^{k - 1} L_km W_mj) (a lane per column).
// Every tile a launch reads was written by an earlier launch; L_ik goes to its
// own array because the step's other blocks still read A_ik.  Then one block
// per lower tile of A_c^-1 (k_cs_wtw) sums W_mi^T W_mj over m >= i and
// writes the FP32 inverse, both triangles.  No pivoting: A_c is symmetric
// positive definite (a non-positive pivot sets *info and the coarse level is
// dropped).  nt + 1 launches of at most bt (bt + 1) / 2 + nt blocks.
// ---------------------------------------------------------------------------
constexpr int kChT = 32;
constexpr int kChTT = kChT * kChT;
constexpr int kChPad = kChT + 1;  // LDS row stride (doubles): rows of a tile fall in distinct banks

struct Band {
  int n;   // order
  int nt;  // tiles per side: ceil(n / T)
  int bt;  // band in tiles
};

__host__ __device__ inline size_t band_slot(const Band &b, int i, int j) { return (size_t)j * (b.bt + 1) + (i - j); }
__host__ __device__ inline size_t tri_slot(int i, int j) { return (size_t)i * (i + 1) / 2 + j; }

// The band tiles of A_c: thread per entry of tiles (i, j), 0 <= i - j <= bt;
// entry (r, c) of A_c sums the coarse elements holding both nodes, in element
// order (zero when they share none); rows or columns past n hold the identity
__global__ void __launch_bounds__(256) k_cs_band(CGrid cg, int dx, int dy, Band b, const double *KE, double *Ab) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (size_t)b.nt * (b.bt + 1) * kChTT) return;
  const int slot = (int)(t / kChTT), e = (int)(t % kChTT);
  const int j = slot / (b.bt + 1), i = j + slot % (b.bt + 1);
  if (i >= b.nt) return;  // past the last tile row
  const int row = i * kChT + e / kChT, col = j * kChT + e % kChT;
  if (row >= b.n || col >= b.n) {
    Ab[t] = row == col ? 1.0 : 0.0;
    return;
  }
  const int qi = row / 3, d = row % 3, qj = col / 3, f = col % 3;
  const int xi = qi % cg.mx + 1, yi = (qi / cg.mx) % cg.my + 1, zi = qi / (cg.mx * cg.my) + 1;
  const int xj = qj % cg.mx + 1, yj = (qj / cg.mx) % cg.my + 1, zj = qj / (cg.mx * cg.my) + 1;
  double s = 0.0;
  if (abs(xi - xj) <= 1 && abs(yi - yj) <= 1 && abs(zi - zj) <= 1)
    for (int ez = max(zi, zj) - 1; ez <= min(zi, zj); ++ez)
      for (int ey = max(yi, yj) - 1; ey <= min(yi, yj); ++ey)
        for (int ex = max(xi, xj) - 1; ex <= min(xi, xj); ++ex) {
          const int A = hex_corner(xi - ex, yi - ey, zi - ez), B = hex_corner(xj - ex, yj - ey, zj - ez);
          s += KE[((size_t)(ez * dy + ey) * dx + ex) * 576 + (A * 3 + d) * 24 + B * 3 + f];
        }
  Ab[t] = s;
}

// A tile in LDS: rows of stride kChPad doubles (kChTile per tile).
constexpr int kChTile = kChT * kChPad;
// tiles of a W row's products loaded together (into registers, then LDS)
constexpr int kChChunk = 4;

// a tile (row-major in global memory) into LDS, by the whole block
__device__ __forceinline__ void tile_in(const double *src, double *dst) {
  for (int e = threadIdx.x; e < kChTT; e += blockDim.x) dst[(e >> 5) * kChPad + (e & 31)] = src[e];
}
__device__ __forceinline__ void tile_out(const double *src, double *dst) {
  for (int e = threadIdx.x; e < kChTT; e += blockDim.x) dst[e] = src[(e >> 5) * kChPad + (e & 31)];
}

typedef double dbl4 __attribute__((ext_vector_type(4)));

// Quadrant (qi, qj) (16 x 16) of a 32 x 32 product C += A B by one wave, with
// FP64 MFMA (16 x 16 x 4 per instruction, p in steps of 4).  a(x, p) =
// A[16 qi + x][p], b(p, x) = B[p][16 qj + x].  The FP64 instruction's
// layouts: the A operand of lane l is row l % 16 at p0 + l / 16, the B
// operand column l % 16 at p0 + l / 16, and accumulator entry v (0 .. 3) of
// lane l is row 4 v + l / 16, column l % 16 of the quadrant (the FP32
// instruction of this shape differs: row 4 (l / 16) + v).
template <class FA, class FB>
__device__ __forceinline__ dbl4 mfma_32(dbl4 c, FA a, FB b) {
  const int l = threadIdx.x & 63, x = l & 15, kk = l >> 4;
#pragma unroll
  for (int p0 = 0; p0 < kChT; p0 += 4) c = __builtin_amdgcn_mfma_f64_16x16x4f64(a(x, p0 + kk), b(p0 + kk, x), c, 0, 0, 0);
  return c;
}
// the tile row / column of accumulator entry v of this lane (wave w: qi = w / 2, qj = w % 2)
__device__ __forceinline__ int acc_row(int v) { return 16 * ((threadIdx.x >> 6) >> 1) + 4 * v + ((threadIdx.x & 63) >> 4); }
__device__ __forceinline__ int acc_col() { return 16 * ((threadIdx.x >> 6) & 1) + (threadIdx.x & 15); }

// 1 / sqrt(x) for a positive normal x: the hardware estimate refined by two
// Newton steps (error within an ulp or so; the factor is a preconditioner's)
__device__ __forceinline__ double rsqrt_nr(double x) {
  const double h = 0.5 * x;
  double y = __builtin_amdgcn_rsq(x);
  y = y * fma(-h * y, y, 1.5);
  return y * fma(-h * y, y, 1.5);
}

__device__ __forceinline__ double lane_bcast(double v, int lane) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                          __builtin_amdgcn_readlane(__double2loint(v), lane));
}

// Cholesky of the 32 x 32 tile t (LDS, lower triangle read) by wave 0: lane r
// holds row r (lanes 32 .. 63 a copy of row r - 32) and keeps its own
// diagonal entry up to date, so the next pivot is read from its lane
// (readlane) as soon as the column's multipliers are known; the multipliers
// go round through LDS (bc, 32 doubles) for the rows' updates, and the next
// pivot's square root is issued while those reads are in flight.  The factor
// goes to dst (row-major, zero above the diagonal); a pivot that is not
// positive sets *info.  The columns are unrolled by templates: as a loop, the
// compiler leaves it partly rolled and moves the row to scratch memory.
template <int J>
__device__ __forceinline__ void potrf_col(double (&a)[kChT], double &dr, double &rs, double &sq, bool &bad, int r,
                                          double *bc) {
  const double l = r > J ? a[J] * rs : (r == J ? sq : 0.0);
  a[J] = l;
  if constexpr (J + 1 < kChT) {
    dr = fma(-l, l, dr);  // = the update of a[r] below, the same operands
    bc[r] = l;    // lanes r and r + 32 store the same value
    __builtin_amdgcn_wave_barrier();
    double lm[kChT - 1 - J];
#pragma unroll
    for (int m = J + 1; m < kChT; ++m) lm[m - J - 1] = bc[m];
    const double piv = lane_bcast(dr, J + 1);
    bad = bad || !(piv > 0.0);
    rs = rsqrt_nr(piv);
    sq = piv * rs;
#pragma unroll
    for (int m = J + 1; m < kChT; ++m) a[m] = fma(-l, lm[m - J - 1], a[m]);
    __builtin_amdgcn_wave_barrier();
  }
}
template <int... J>
__device__ __forceinline__ void potrf_cols(double (&a)[kChT], double &dr, double &rs, double &sq, bool &bad, int r,
                                           double *bc, std::integer_sequence<int, J...>) {
  (potrf_col<J>(a, dr, rs, sq, bad, r, bc), ...);
}
__device__ __forceinline__ void tile_potrf(const double *t, double *dst, int *info, double *bc) {
  const int lane = threadIdx.x, r = lane & (kChT - 1);
  double a[kChT];
#pragma unroll
  for (int c = 0; c < kChT; ++c) a[c] = t[r * kChPad + c];
  double dr = a[0];
#pragma unroll
  for (int c = 1; c < kChT; ++c) dr = r == c ? a[c] : dr;
  const double piv = t[0];
  bool bad = !(piv > 0.0);
  double rs = rsqrt_nr(piv), sq = piv * rs;
  potrf_cols(a, dr, rs, sq, bad, r, bc, std::make_integer_sequence<int, kChT>{});
  if (lane < kChT)
#pragma unroll
    for (int c = 0; c < kChT; ++c) dst[r * kChT + c] = c <= r ? a[c] : 0.0;
  if (bad && lane == 0) *info = 1;
}

// x = L^-1 x for the lower triangular L in LDS (rows of stride kChPad) with
// dinv = 1 / diag(L), column by column (after x[c] is final, the later
// entries take its term), so one multiply-add per column is on the chain
__device__ __forceinline__ void trsv_lower(double (&x)[kChT], const double *L, const double *dinv) {
#pragma unroll
  for (int c = 0; c < kChT; ++c) {
    x[c] *= dinv[c];
#pragma unroll
    for (int m = c + 1; m < kChT; ++m) x[m] = fma(-L[m * kChPad + c], x[c], x[m]);
  }
}

__global__ void __launch_bounds__(256) k_cs_chol(Band b, int k, double *Ab, double *Lb, double *Wp, int *info) {
  // LDS: L_kk | the tile for W's solve or the factor | 1 / diag(L_kk) | the
  // factor's exchange | A_ik, A_jk (step) or the chunk's tiles (W row)
  __shared__ double smem[(2 + 2 * kChChunk) * kChTile + 2 * kChT];
  double *sL = smem, *sF = smem + kChTile, *sD = smem + 2 * kChTile, *bc = sD + kChT, *sX = bc + kChT;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int qi = wave >> 1, qj = wave & 1, ac = acc_col();
  if (k < 0) {  // factor tile (0, 0)
    tile_in(Ab, sF);
    __syncthreads();
    if (wave == 0) tile_potrf(sF, Lb, info, bc);
    return;
  }
  const int m = min(b.bt, b.nt - 1 - k), nch = m * (m + 1) / 2;
  const double *lkk = Lb + band_slot(b, k, k) * kChTT;
  tile_in(lkk, sL);
  if (tid < kChT) sD[tid] = 1.0 / lkk[tid * (kChT + 1)];
  if ((int)blockIdx.x < nch) {
    // ---- the right-looking step: tile (i, j) ----
    int di = 0, rest = blockIdx.x;
    while (rest > di) {
      rest -= di + 1;
      ++di;
    }
    const int i = k + 1 + di, j = k + 1 + rest;
    double *sXi = sX, *sXj = i != j ? sX + kChTile : sX;
    tile_in(Ab + band_slot(b, i, k) * kChTT, sXi);
    if (i != j) tile_in(Ab + band_slot(b, j, k) * kChTT, sXj);
    __syncthreads();
    // L_ik = A_ik L_kk^-T (lanes 0..31) and L_jk (lanes 32..63): row r solves
    // L_kk x = a_r
    if (wave == 0 && (lane < kChT || i != j)) {
      double *t = lane < kChT ? sXi : sXj;
      const int r = lane & (kChT - 1);
      double x[kChT];
#pragma unroll
      for (int c = 0; c < kChT; ++c) x[c] = t[r * kChPad + c];
      trsv_lower(x, sL, sD);
#pragma unroll
      for (int c = 0; c < kChT; ++c) t[r * kChPad + c] = x[c];
    }
    __syncthreads();
    // A_ij -= L_ik L_jk^T, a quadrant per wave
    double *aij = Ab + band_slot(b, i, j) * kChTT;
    dbl4 acc;
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[v] = aij[acc_row(v) * kChT + ac];
    acc = mfma_32(
        acc, [&](int x, int p) { return sXi[(16 * qi + x) * kChPad + p]; },
        [&](int p, int x) { return -sXj[(16 * qj + x) * kChPad + p]; });
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      aij[acc_row(v) * kChT + ac] = acc[v];
      sF[acc_row(v) * kChPad + ac] = acc[v];
    }
    if (j == k + 1) tile_out(sXi, Lb + band_slot(b, i, k) * kChTT);
    if (i == k + 1 && j == k + 1) {
      __syncthreads();
      if (wave == 0) tile_potrf(sF, Lb + band_slot(b, k + 1, k + 1) * kChTT, info, bc);
    }
    return;
  }
  // ---- row k of W: tile (k, j) ----
  const int j = blockIdx.x - nch;
  dbl4 acc;
#pragma unroll
  for (int v = 0; v < 4; ++v) acc[v] = (j == k && acc_row(v) == ac) ? 1.0 : 0.0;
  // chunks of kChChunk products: the chunk's tiles L_k,mm and W_mm,j load
  // into registers (all loads issued before any store) while the previous
  // chunk multiplies, then go to LDS
  const int cbeg = max(j, k - b.bt);
  double rg[2 * kChChunk][4];
  auto load_chunk = [&](int c0) {
#pragma unroll
    for (int t = 0; t < 2 * kChChunk; ++t) {
      const int mm = c0 + t / 2;
      const double *src = mm >= k ? Lb : (t & 1) ? Wp + tri_slot(mm, j) * kChTT : Lb + band_slot(b, k, mm) * kChTT;
#pragma unroll
      for (int s = 0; s < 4; ++s) rg[t][s] = src[tid + 256 * s];
    }
  };
  if (cbeg < k) load_chunk(cbeg);
  for (int c0 = cbeg; c0 < k; c0 += kChChunk) {
    __syncthreads();  // the previous chunk's tiles are read
#pragma unroll
    for (int t = 0; t < 2 * kChChunk; ++t)
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int e = tid + 256 * s;
        sX[t * kChTile + (e >> 5) * kChPad + (e & 31)] = rg[t][s];
      }
    __syncthreads();
    if (c0 + kChChunk < k) load_chunk(c0 + kChChunk);
#pragma unroll
    for (int t = 0; t < kChChunk; ++t) {
      if (c0 + t >= k) break;
      const double *lt = sX + 2 * t * kChTile, *wt = lt + kChTile;
      acc = mfma_32(
          acc, [&](int x, int p) { return lt[(16 * qi + x) * kChPad + p]; },
          [&](int p, int x) { return -wt[p * kChPad + 16 * qj + x]; });
    }
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) sF[acc_row(v) * kChPad + ac] = acc[v];
  __syncthreads();
  // W_kj = L_kk^-1 R: lane c solves column c
  if (wave == 0 && lane < kChT) {
    const int c = lane;
    double w[kChT];
#pragma unroll
    for (int r = 0; r < kChT; ++r) w[r] = sF[r * kChPad + c];
    trsv_lower(w, sL, sD);
    double *dst = Wp + tri_slot(k, j) * kChTT;
#pragma unroll
    for (int r = 0; r < kChT; ++r) dst[r * kChT + c] = w[r];
  }
}

// A_c^-1 = W^T W: block per lower tile (i, j) (tri_slot order), tile (i, j)
// = sum over m = i .. nt - 1 of W_mi^T W_mj (FP64 MFMA, a quadrant per wave;
// the next m's tiles load into registers while the current ones multiply, two
// LDS buffers).  The FP32 result goes to ainv (n x n, symmetric) at (i, j)
// and, through LDS, transposed at (j, i); a diagonal tile is written from its
// lower triangle, so the inverse is exactly symmetric.
__global__ void __launch_bounds__(256) k_cs_wtw(Band b, const double *Wp, float *ainv) {
  __shared__ double smem[4 * kChTile];
  int i = 0, rest = blockIdx.x;
  while (rest > i) {
    rest -= i + 1;
    ++i;
  }
  const int j = rest, tid = threadIdx.x, wave = tid >> 6, qi = wave >> 1, qj = wave & 1, ac = acc_col();
  dbl4 acc = {0.0, 0.0, 0.0, 0.0};
  double rg[2][4];
  auto load = [&](int m) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      rg[0][s] = Wp[tri_slot(m, i) * kChTT + tid + 256 * s];
      rg[1][s] = Wp[tri_slot(m, j) * kChTT + tid + 256 * s];
    }
  };
  load(i);
  for (int m = i; m < b.nt; ++m) {
    double *wi = smem + (m & 1) * 2 * kChTile, *wj = wi + kChTile;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int e = tid + 256 * s, o = (e >> 5) * kChPad + (e & 31);
      wi[o] = rg[0][s];
      wj[o] = rg[1][s];
    }
    __syncthreads();  // (the buffer's previous readers passed the last barrier)
    if (m + 1 < b.nt) load(m + 1);
    acc = mfma_32(
        acc, [&](int x, int p) { return wi[p * kChPad + 16 * qi + x]; },
        [&](int p, int x) { return wj[p * kChPad + 16 * qj + x]; });
  }
  __syncthreads();
  double *o = smem;
#pragma unroll
  for (int v = 0; v < 4; ++v) o[acc_row(v) * kChPad + ac] = acc[v];
  __syncthreads();
  const int n = b.n;
  for (int e = tid; e < kChTT; e += 256) {
    const int r = e >> 5, c = e & 31;
    const int row = i * kChT + r, col = j * kChT + c;
    const double v = i == j ? o[max(r, c) * kChPad + min(r, c)] : o[r * kChPad + c];
    if (row < n && col < n) ainv[(size_t)row * n + col] = (float)v;
  }
  if (i == j) return;
  for (int e = tid; e < kChTT; e += 256) {
    const int r = e >> 5, c = e & 31;  // entry (c, r) of tile (i, j) at (j T + r, i T + c)
    const int row = j * kChT + r, col = i * kChT + c;
    if (row < n && col < n) ainv[(size_t)row * n + col] = (float)o[c * kChPad + r];
  }
}

}  // namespace

void coarse_free(CoarseDev *c) {
  if (!c) return;
  dfree(c->arena);
  delete c;
}

bool coarse_enabled() {
  const char *env = std::getenv("MPPD_COARSE");
  return !(env && std::atoi(env) == 0);
}

// The coarsest-first search: the largest coarse grid aligned with g (a
// divisor of each side's element count, at least 2 coarse elements a side)
// with at most max_dofs coarse DOFs and at most kCoarseMaxPos fine elements
// per coarse element.
bool coarse_grid(const Grid &g, int max_dofs, CGrid *out) {
  if (g.nx > kCoarseMaxLine || g.ny > kCoarseMaxLine || g.nz > kCoarseMaxLine) return false;
  int dx = 0, dy = 0, dz = 0;
  for (int dmax = 32; dmax >= 2; --dmax) {
    dx = pick_side(g.nex, dmax);
    dy = pick_side(g.ney, dmax);
    dz = pick_side(g.nez, dmax);
    if (dx && dy && dz && 3 * (dx - 1) * (dy - 1) * (dz - 1) <= max_dofs) break;
    dx = dy = dz = 0;
  }
  if (!dx) return false;
  CGrid cg;
  cg.hx = g.nex / dx;
  cg.hy = g.ney / dy;
  cg.hz = g.nez / dz;
  cg.mx = dx - 1;
  cg.my = dy - 1;
  cg.mz = dz - 1;
  cg.nq = cg.mx * cg.my * cg.mz;
  cg.nc = 3 * cg.nq;
  if (cg.hx * cg.hy * cg.hz > kCoarseMaxPos) return false;
  *out = cg;
  return true;
}

int coarse_inverse(const Grid &g, const CGrid &cg, const int *et_d, const double *ke_d, float *ainv, hipStream_t st,
                   bool *ok) {
  *ok = false;
  const int nc = cg.nc, npos = cg.hx * cg.hy * cg.hz;
  const int dx = cg.mx + 1, dy = cg.my + 1, ncel = dx * dy * (cg.mz + 1);
  Band b;
  b.n = nc;
  b.nt = (nc + kChT - 1) / kChT;
  b.bt = (3 * (cg.mx * cg.my + cg.mx + 1) + 2 + kChT - 1) / kChT;
  const size_t nband = (size_t)b.nt * (b.bt + 1) * kChTT, nw = tri_slot(b.nt, 0) * kChTT;
  const size_t nC = (size_t)npos * 3 * 576, nKE = (size_t)ncel * 576;
  double *buf = nullptr;
  int *info_d = nullptr;
  int rc = MPPD_OK;
  do {
    // one allocation for the temporaries: Ab | Lb | Wp | C | KE
    if ((rc = dalloc(&buf, 2 * nband + nw + nC + nKE, false))) break;
    if ((rc = dalloc(&info_d, 1, false))) break;
    double *Ab = buf, *Lb = Ab + nband, *Wp = Lb + nband, *C = Wp + nw, *KE = C + nC;
    hipError_t e = hipMemsetAsync(info_d, 0, sizeof(int), st);
    if (e != hipSuccess) {
      rc = hip_fail(e, "coarse factor status");
      break;
    }
    hipLaunchKernelGGL(k_cs_table, dim3(npos, 3), dim3(576), 0, st, cg, ke_d, C);
    hipLaunchKernelGGL(k_cs_elem, dim3(ncel), dim3(576), 0, st, g, cg, dx, dy, et_d, C, KE);
    hipLaunchKernelGGL(k_cs_band, dim3((unsigned)((nband + 255) / 256)), dim3(256), 0, st, cg, dx, dy, b, KE, Ab);
    for (int k = -1; k < b.nt; ++k) {
      const int m = k < 0 ? 0 : std::min(b.bt, b.nt - 1 - k);
      const int nblk = k < 0 ? 1 : m * (m + 1) / 2 + k + 1;
      hipLaunchKernelGGL(k_cs_chol, dim3(nblk), dim3(256), 0, st, b, k, Ab, Lb, Wp, info_d);
    }
    hipLaunchKernelGGL(k_cs_wtw, dim3((unsigned)tri_slot(b.nt, 0)), dim3(256), 0, st, b, Wp, ainv);
    int info = 0;
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&info, info_d, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      rc = hip_fail(e, "coarse matrix factorisation");
      break;
    }
    *ok = info == 0;  // else not positive definite: no coarse level
  } while (0);
  (void)hipStreamSynchronize(st);  // the temporaries below may still be in use after a failed step
  dfree(buf);
  dfree(info_d);
  return rc;
}

// Builds the coarse level of a K-field system on the RVE grid g whose phase
// per element is et_d and whose per-phase element stiffness is ke_d (3 x 576,
// FP64, on the device).  *out = nullptr (Jacobi alone) when the grid has no
// usable coarse grid (a side with fewer than 4 elements or prime, a side
// beyond kCoarseMaxLine nodes), when A_c is not positive definite, or when
// MPPD_COARSE=0.
int coarse_setup(const Grid &g, const int *et_d, const double *ke_d, int K, hipStream_t st, CoarseDev **out) {
  *out = nullptr;
  CGrid cg;
  if (!coarse_enabled() || !coarse_grid(g, kCoarseMaxDofs, &cg)) return MPPD_OK;
  const int nc = cg.nc;
  CoarseDev *c = new CoarseDev();
  c->cg = cg;
  int rc = MPPD_OK;
  bool ok = false;
  do {
    // one allocation, 256-byte aligned parts: the inverse (written whole by
    // coarse_inverse) last, everything before it zero-filled on st in one
    // fill (coarse_inverse synchronizes st)
    size_t off = 0;
    auto carve = [&off](size_t bytes) {
      const size_t o = off;
      off += (bytes + 255) & ~(size_t)255;
      return o;
    };
    const size_t o_t1 = carve((size_t)3 * K * g.nz * g.ny * cg.mx * sizeof(float));
    const size_t o_rc = carve((size_t)nc * K * sizeof(float)), o_ecp = carve((size_t)kCoarseSplit * nc * K * sizeof(float));
    const size_t o_ec = carve((size_t)nc * K * sizeof(float)), o_z = carve(mixed_vector_len(g, K) * sizeof(float));
    const size_t o_prz = carve((size_t)3 * K * coarse_vec_blocks(g) * sizeof(double));
    const size_t zeroed = off, o_ainv = carve((size_t)nc * nc * sizeof(float));
    hipError_t e = hipMalloc(&c->arena, off);
    if (e == hipSuccess) e = hipMemsetAsync(c->arena, 0, zeroed, st);
    if (e != hipSuccess) {
      rc = hip_fail(e, "coarse buffers");
      break;
    }
    char *base = static_cast<char *>(c->arena);
    c->t1 = reinterpret_cast<float *>(base + o_t1);
    c->rc = reinterpret_cast<float *>(base + o_rc);
    c->ecp = reinterpret_cast<float *>(base + o_ecp);
    c->ec = reinterpret_cast<float *>(base + o_ec);
    c->z = reinterpret_cast<float *>(base + o_z);
    c->prz = reinterpret_cast<double *>(base + o_prz);
    float *ainv = reinterpret_cast<float *>(base + o_ainv);
    c->ainv = ainv;
    rc = coarse_inverse(g, cg, et_d, ke_d, ainv, st, &ok);
  } while (0);
  if (rc || !ok) {
    (void)hipStreamSynchronize(st);  // the zero fills of an early failure may still be queued
    coarse_free(c);
    return rc;
  }
  *out = c;
  return MPPD_OK;
}

}  // namespace mpp
