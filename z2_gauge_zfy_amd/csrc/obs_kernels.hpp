// Equal-time observables of one slice (reference Obs_equal_calc,
// src/obser_equal.f90:73-149) accumulated per chain on the device, and the
// per-bin averages (Obs_equal_ave, :50-71).
#pragma once
#include "z2_common.hpp"

// per-chain accumulator layout (doubles)
enum : int {
  OBS_FLUX_SUM = 0, OBS_FLUX_SQ, OBS_PAIR, OBS_CDW, OBS_CDW_SHIFT, OBS_DENSITY, OBS_NOBS,
  OBS_NACC  // 7 accumulators; results below
};
enum : int {
  RES_FLUX_AVG = 0, RES_CHI_FLUX, RES_PAIR_SC_Q0, RES_CDW_PI, RES_CDW_SHIFT, RES_CDW_RATIO,
  RES_DENSITY, RES_N
};

template <typename R>
struct ObsArgs {
  const R *G;
  const int8_t *sig;
  int T, nt;
  double *acc;           // chain stride OBS_NACC
  const double *cosT;    // (2Lx-1) x (2Ly-1) table of cos(qshift_x*dx + pi*dy)
  int Lx, Ly;
  LatticeDev lat;
  double *part;          // per chain and slot: part_stride x 3 tile partial sums
  int part_stride;
  double *diag;          // per chain and slot: 1 - G(i, i), N values
  double *flux;          // per chain and slot: the plaquette products of the slice, N values
  int slot;              // k_obs_tiles: the slot this slice's values go to
};

// slots per chain: the slices whose sums wait for k_obs_finish
constexpr int OBS_SLOTS = 16;

// The N x N pair sums (pair_sc_q0, cdw_pi, cdw_shift) of one slice, one 32 x 32
// tile of (i, j) pairs per workgroup: G(i, j) and G(j, i) of the tile are staged
// in LDS by coalesced column reads; the tile's three partial sums go to slot
// a.slot of `part`, and k_obs_finish adds them by a fixed tree (deterministic).
// The diagonal tiles also store 1 - G(i, i) of their rows and the first tile the
// slice's plaquette products, so that nothing k_obs_finish reads changes before
// it runs (the flips that follow change G and the slice's field).
#define OBS_TB 32
template <typename R>
__global__ void __launch_bounds__(256) k_obs_tiles(ObsArgs<R> a) {
  const int N = a.lat.N;
  const int chain = blockIdx.y, nt = (N + OBS_TB - 1) / OBS_TB;
  const int bi = blockIdx.x % nt, bj = blockIdx.x / nt;
  const int i0 = bi * OBS_TB, j0 = bj * OBS_TB;
  const R *G = a.G + (size_t)chain * N * N;
  __shared__ double A[OBS_TB][OBS_TB + 1];  // A[v][u] = G(i0 + u, j0 + v)
  __shared__ double B[OBS_TB][OBS_TB + 1];  // B[u][v] = G(j0 + v, i0 + u)
  __shared__ double ni[OBS_TB], nj[OBS_TB];
  __shared__ double red[48];
  const int tid = threadIdx.x;
  for (int e = tid; e < OBS_TB * OBS_TB; e += blockDim.x) {
    const int c = e / OBS_TB, r = e % OBS_TB;  // constant divisor
    const int ia = i0 + r, ja = j0 + c, ib = j0 + r, jb = i0 + c;
    A[c][r] = (ia < N && ja < N) ? (double)G[ia + (size_t)ja * N] : 0.0;
    B[c][r] = (ib < N && jb < N) ? (double)G[ib + (size_t)jb * N] : 0.0;
  }
  const size_t cs = (size_t)chain * OBS_SLOTS + a.slot;  // (chain, slot)
  if (tid < OBS_TB) {
    const int i = i0 + tid;
    ni[tid] = i < N ? 1.0 - (double)G[i + (size_t)i * N] : 0.0;
    if (bi == bj && i < N) a.diag[cs * N + i] = ni[tid];
  } else if (tid < 2 * OBS_TB) {
    const int j = j0 + tid - OBS_TB;
    nj[tid - OBS_TB] = j < N ? 1.0 - (double)G[j + (size_t)j * N] : 0.0;
  }
  if (blockIdx.x == 0) {
    const int8_t *s = a.sig + (size_t)chain * a.T * a.lat.N2 + (size_t)(a.nt - 1) * a.lat.N2;
    for (int i = tid; i < N; i += blockDim.x) {
      const int *pl = a.lat.plaq + 4 * i;
      a.flux[cs * N + i] = (double)((int)s[pl[0]] * (int)s[pl[1]] * (int)s[pl[2]] * (int)s[pl[3]]);
    }
  }
  __syncthreads();
  double pair = 0.0, cdw = 0.0, cdws = 0.0;
  const int W = 2 * a.Lx - 1;
  for (int e = tid; e < OBS_TB * OBS_TB; e += blockDim.x) {
    const int v = e / OBS_TB, u = e % OBS_TB;
    const int ii = i0 + u, jj = j0 + v;
    if (ii >= N || jj >= N) continue;
    const double n_i = ni[u], n_j = nj[v];
    const double delta = ii == jj ? 1.0 : 0.0;
    const double gij = A[v][u], gji = B[u][v];
    const double x = delta - gij;
    pair += x * x;
    double dc = n_i * n_j - (delta - gji) * gij;
    dc = dc + n_i * n_j - (delta - gji) * gij;
    dc = dc + n_i * n_j + n_i * n_j;
    const double ph = (double)(a.lat.stag[ii] * a.lat.stag[jj]);
    const int dx = a.lat.xy[2 * ii] - a.lat.xy[2 * jj] + a.Lx - 1;
    const int dy = a.lat.xy[2 * ii + 1] - a.lat.xy[2 * jj + 1] + a.Ly - 1;
    cdw += ph * dc;
    cdws += a.cosT[dx + W * dy] * dc;
  }
  double sums[3] = {pair, cdw, cdws};
  block_reduce_sum_n<double, 3>(sums, red);
  if (tid == 0) {
    double *p = a.part + (cs * a.part_stride + blockIdx.x) * 3;
    p[0] = sums[0];
    p[1] = sums[1];
    p[2] = sums[2];
  }
}

// density and flux of slots 0 .. nslots-1 (slices in the order they were
// measured), the sums of their tile partials, and the accumulation into the bin
// slice by slice (one workgroup of 256 threads per chain): every slice's sums and
// additions are those of one launch per slice
__global__ void __launch_bounds__(256) k_obs_finish(ObsArgs<double> a, int nslots) {
  const int N = a.lat.N;
  const int chain = blockIdx.x, ntile = ((N + OBS_TB - 1) / OBS_TB) * ((N + OBS_TB - 1) / OBS_TB);
  __shared__ double red[80];
  for (int sl = 0; sl < nslots; ++sl) {
    const size_t cs = (size_t)chain * OBS_SLOTS + sl;
    const double *dg = a.diag + cs * N, *fx = a.flux + cs * N;
    double dens = 0.0, flux = 0.0;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
      const double nd = dg[i];
      dens += nd + nd;
      flux += fx[i];
    }
    const double *p = a.part + cs * a.part_stride * 3;
    double pair = 0.0, cdw = 0.0, cdws = 0.0;
    for (int t = threadIdx.x; t < ntile; t += blockDim.x) {
      pair += p[3 * t];
      cdw += p[3 * t + 1];
      cdws += p[3 * t + 2];
    }
    double sums[5] = {dens, flux, pair, cdw, cdws};
    block_reduce_sum_n<double, 5>(sums, red);
    dens = sums[0];
    flux = sums[1];
    pair = sums[2];
    cdw = sums[3];
    cdws = sums[4];
    if (threadIdx.x == 0) {
      double *acc = a.acc + (size_t)chain * OBS_NACC;
      const double ft = flux / (double)N;
      acc[OBS_FLUX_SUM] += ft;
      acc[OBS_FLUX_SQ] += ft * ft;
      acc[OBS_PAIR] += pair;
      acc[OBS_CDW] += cdw;
      acc[OBS_CDW_SHIFT] += cdws;
      acc[OBS_DENSITY] += dens / (double)N;
      acc[OBS_NOBS] += 1.0;
    }
  }
}

// per-bin averages (Obs_equal_ave)
__global__ void k_obs_ave(const double *acc, double *res, const ChainConst *cc, int N, int nch, int nobs_in) {
  int chain = blockIdx.x * blockDim.x + threadIdx.x;
  if (chain >= nch) return;
  const double *a = acc + (size_t)chain * OBS_NACC;
  double *r = res + (size_t)chain * RES_N;
  double nobs = nobs_in > 0 ? (double)nobs_in : a[OBS_NOBS];
  if (nobs <= 0.0) return;
  double z = 1.0 / nobs;
  double fm = a[OBS_FLUX_SUM] * z, fsq = a[OBS_FLUX_SQ] * z;
  r[RES_FLUX_AVG] = fm;
  double v = fsq - fm * fm;
  r[RES_CHI_FLUX] = cc[chain].beta * (double)N * (v > 0.0 ? v : 0.0);
  r[RES_PAIR_SC_Q0] = (a[OBS_PAIR] * z) / (double)N;
  r[RES_CDW_PI] = (a[OBS_CDW] * z) / (double)N;
  r[RES_CDW_SHIFT] = (a[OBS_CDW_SHIFT] * z) / (double)N;
  r[RES_CDW_RATIO] = r[RES_CDW_PI] > 1e-12 ? 1.0 - r[RES_CDW_SHIFT] / r[RES_CDW_PI] : 0.0;
  r[RES_DENSITY] = a[OBS_DENSITY] * z;
}
