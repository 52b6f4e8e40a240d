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
};

// One workgroup per chain.
template <typename R>
__global__ void __launch_bounds__(1024) k_obs_equal(ObsArgs<R> a) {
  const int N = a.lat.N, N2 = a.lat.N2;
  const int chain = blockIdx.x;
  const R *G = a.G + (size_t)chain * N * N;
  const int8_t *s = a.sig + (size_t)chain * a.T * N2 + (size_t)(a.nt - 1) * N2;
  extern __shared__ unsigned char smem_raw[];
  double *nd = reinterpret_cast<double *>(smem_raw);  // N: n_i = 1 - G_ii
  __shared__ double red[16];
  for (int i = threadIdx.x; i < N; i += blockDim.x) nd[i] = 1.0 - (double)G[i + (size_t)i * N];
  __syncthreads();
  double pair = 0.0, cdw = 0.0, cdws = 0.0, dens = 0.0, flux = 0.0;
  const int W = 2 * a.Lx - 1;
  for (int idx = threadIdx.x; idx < N * N; idx += blockDim.x) {
    int jj = idx / N, ii = idx - jj * N;  // consecutive threads: consecutive ii (coalesced G(ii,jj))
    double n_i = nd[ii], n_j = nd[jj];
    double delta = ii == jj ? 1.0 : 0.0;
    double gij = (double)G[ii + (size_t)jj * N], gji = (double)G[jj + (size_t)ii * N];
    double x = delta - gij;
    pair += x * x;
    double dc = n_i * n_j - (delta - gji) * gij;
    dc = dc + n_i * n_j - (delta - gji) * gij;
    dc = dc + n_i * n_j + n_i * n_j;
    double ph = (double)(a.lat.stag[ii] * a.lat.stag[jj]);
    int dx = a.lat.xy[2 * ii] - a.lat.xy[2 * jj] + a.Lx - 1;
    int dy = a.lat.xy[2 * ii + 1] - a.lat.xy[2 * jj + 1] + a.Ly - 1;
    cdw += ph * dc;
    cdws += a.cosT[dx + W * dy] * dc;
  }
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    dens += nd[i] + nd[i];
    const int *pl = a.lat.plaq + 4 * i;
    flux += (double)((int)s[pl[0]] * (int)s[pl[1]] * (int)s[pl[2]] * (int)s[pl[3]]);
  }
  pair = block_reduce_sum(pair, red);
  cdw = block_reduce_sum(cdw, red);
  cdws = block_reduce_sum(cdws, red);
  dens = block_reduce_sum(dens, red);
  flux = block_reduce_sum(flux, red);
  if (threadIdx.x == 0) {
    double *acc = a.acc + (size_t)chain * OBS_NACC;
    double ft = flux / (double)N;
    acc[OBS_FLUX_SUM] += ft;
    acc[OBS_FLUX_SQ] += ft * ft;
    acc[OBS_PAIR] += pair;
    acc[OBS_CDW] += cdw;
    acc[OBS_CDW_SHIFT] += cdws;
    acc[OBS_DENSITY] += dens / (double)N;
    acc[OBS_NOBS] += 1.0;
  }
}

// per-bin averages (Obs_equal_ave)
__global__ void k_obs_ave(const double *acc, double *res, const ChainConst *cc, int N, int nch) {
  int chain = blockIdx.x * blockDim.x + threadIdx.x;
  if (chain >= nch) return;
  const double *a = acc + (size_t)chain * OBS_NACC;
  double *r = res + (size_t)chain * RES_N;
  double nobs = a[OBS_NOBS];
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
