// Common definitions for the gfx950 Z2-gauge DQMC kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define Z2_WAVE 64

// Per-chain constants computed on the host with the same libm calls the
// reference makes (cosh/sinh/exp/tanh/log), so device arithmetic that uses them
// reproduces the reference values exactly.
struct ChainConst {
  double Cp, Cm, Sp, Sm;  // cosh(+a), cosh(-a), sinh(+a), sinh(-a), a = dtau*t/2
  double delta[2][4];     // Delta for s_old = +1 ([0]) and s_old = -1 ([1]), col-major 2x2
  double boson[9];        // exp(alpha*dS) [* exp(gamma*dT)], index (dS/4+1)*3 + (dT/4+1)
  double ep, em;          // exp(+dtau*mu), exp(-dtau*mu)
  double beta;
  double rt;              // hopping t (current-current correlator prefactor)
  double ptime;           // Wolff: probability to bind equal neighbouring time links
  double walpha;          // Wolff: dtau * J of the plaquette log ratio
};

// lattice tables shared by all chains of one context (same L)
struct LatticeDev {
  int N, N2;
  int Lx, Ly;
  const int *bond_p;       // 2N x 2 sites (0-based): [b*2+k]
  const int *grp;          // 4 groups x (N/2) bonds (0-based), group g at grp + g*(N/2)
  const int *bond_bonds;   // 2N x 6 neighbour bonds (0-based): [b*6+k]
  const int *stag;         // N: +1/-1 sublattice sign
  const int *xy;           // N x 2 coordinates (1-based)
  const int *plaq;         // N x 4 bonds of the plaquette at each site (bottom,right,top,left)
};

__device__ __forceinline__ int kin_group_of_stage(int s) { return s < 4 ? s : 7 - s; }

// C sums at once, each by the tree of block_reduce_sum (bit-identical to C calls of
// it); sh holds 16 C values
template <typename T, int C>
__device__ __forceinline__ void block_reduce_sum_n(T (&v)[C], T *sh) {
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int q = 0; q < C; ++q) v[q] += __shfl_down(v[q], off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < C; ++q) sh[16 * q + wid] = v[q];
  __syncthreads();
  const int nw = (blockDim.x + 63) >> 6;
  if (threadIdx.x < C) {
    T r = 0;
    for (int w = 0; w < nw; ++w) r += sh[16 * threadIdx.x + w];
    sh[16 * threadIdx.x] = r;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < C; ++q) v[q] = sh[16 * q];
  __syncthreads();
}

template <typename T>
__device__ __forceinline__ T block_reduce_sum(T v, T *sh) {
  // 64-wide wavefront reduction then across waves via LDS
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  int nw = (blockDim.x + 63) >> 6;
  T r = 0;
  if (threadIdx.x == 0) {
    for (int w = 0; w < nw; ++w) r += sh[w];
    sh[0] = r;
  }
  __syncthreads();
  r = sh[0];
  __syncthreads();
  return r;
}
