// reduce.hpp -- deterministic block and grid reductions for gfx950.
//
// A grid reduction is finished by the last block to arrive: every block
// writes its partial, counts itself in with an agent-scope atomic, and the
// block that sees the count reach gridDim.x sums the partials in index order
// and resets the counter.  The tree shapes are fixed, so results are
// reproducible run to run, and no host synchronisation or memset is needed
// (the pattern is safe inside a captured graph).
#pragma once

#include <hip/hip_runtime.h>

namespace mpp {

constexpr int kWave = 64;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Wave sum through DPP lane moves instead of LDS-routed shuffles (about 3x
// less latency).  The sequence: quad_perm [1,0,3,2] and [2,3,0,1], row_ror 4
// and 8 (every lane holds its 16-lane row sum), row_bcast 15 and 31 (lane 63
// accumulates the four rows); lane 63's value is then read into every lane.
// Lanes without a DPP source add 0 (update_dpp with old = 0).
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double wave_sum_dpp(double v) {
  v += dpp_move<0xb1>(v);
  v += dpp_move<0x4e>(v);
  v += dpp_move<0x124>(v);
  v += dpp_move<0x128>(v);
  v += dpp_move<0x142>(v);
  v += dpp_move<0x143>(v);
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
  return __hiloint2double(hi, lo);
}

// the same DPP sequence for a float (one 32-bit move per step)
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}

__device__ __forceinline__ float wave_sum_dpp(float v) {
  v += dpp_move<0xb1>(v);
  v += dpp_move<0x4e>(v);
  v += dpp_move<0x124>(v);
  v += dpp_move<0x128>(v);
  v += dpp_move<0x142>(v);
  v += dpp_move<0x143>(v);
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Sum NV values over a block of BS threads; the totals are returned to every
// thread.  lds must hold (BS/64)*NV elements.
template <int NV, int BS, typename T>
__device__ __forceinline__ void block_sum(T (&v)[NV], T *lds) {
  const int tid = threadIdx.x, lane = tid & (kWave - 1), w = tid / kWave;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
  __syncthreads();  // lds may still be read by a previous use
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) lds[w * NV + k] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    T s = 0;
    for (int i = 0; i < BS / kWave; ++i) s += lds[i * NV + k];
    v[k] = s;
  }
}

// Grid-wide sum.  Returns true in (every thread of) the last block, where
// out[] holds the totals; other blocks return false.  partials needs
// gridDim.x * NV elements, *counter must be 0 on entry and is 0 on exit.
template <int NV, int BS, typename T>
__device__ bool grid_sum(T (&v)[NV], T *partials, unsigned *counter, T (&out)[NV], T *lds) {
  __shared__ int last;
  block_sum<NV, BS>(v, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[(size_t)blockIdx.x * NV + k] = v[k];
    __threadfence();
    const unsigned prev = atomicAdd(counter, 1u);
    last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return false;
  __threadfence();
  T acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0;
  for (unsigned b = threadIdx.x; b < gridDim.x; b += BS)
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] += partials[(size_t)b * NV + k];
  block_sum<NV, BS>(acc, lds);
#pragma unroll
  for (int k = 0; k < NV; ++k) out[k] = acc[k];
  if (threadIdx.x == 0) *counter = 0u;
  return true;
}

}  // namespace mpp
