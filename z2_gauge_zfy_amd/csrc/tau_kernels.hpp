// Time-displaced Green's functions and observables (reference Dyn%sweep_R,
// src/dynamics.f90:41-62; Wrap_tau / stab_green_big, src/stabilization.f90:
// 202-264, 478-514; Obs_tau_calc, src/obser_tau.f90:56-133).  FP64.
#pragma once
#include "z2_common.hpp"

enum : int { TAU_SPIN = 0, TAU_PAIR_S, TAU_PAIR_D, TAU_CHARGE_S, TAU_CHARGE_D, TAU_SINGLE, TAU_CURXX,
             TAU_NOBS };

// Dyn_reset: Gr00 = Grtt = Grt0 = G, Gr0t = -(I - G)  (src/process_matrix.f90:88-96)
template <typename R>
__global__ void __launch_bounds__(256) k_dyn_reset(const R *G, double *G00, double *Gtt, double *Gt0,
                                                   double *G0t, int N) {
  const int chain = blockIdx.y;
  const size_t off = (size_t)chain * N * N;
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < N * N; idx += gridDim.x * blockDim.x) {
    const int c = idx / N, r = idx - c * N;
    const double g = (double)G[off + idx];
    G00[off + idx] = g;
    Gtt[off + idx] = g;
    Gt0[off + idx] = g;
    G0t[off + idx] = -((r == c ? 1.0 : 0.0) - g);
  }
}

// Obs_tau_calc for slice nt, both spins equal, real parts.  One workgroup per
// chain, one thread per displacement bin d = imj(i,j) - 1; the bin is
// accumulated over j in increasing order, as the reference's jj loop does.
struct TauObsArgs {
  const double *G00, *Gtt, *Gt0, *G0t;
  const int8_t *sig;
  double *acc;  // chain stride 7*T*N: [k][nt-1][d]
  const ChainConst *cc;
  int T, nt, Lx, Ly;
  LatticeDev lat;
};

__global__ void __launch_bounds__(1024) k_obs_tau(TauObsArgs a) {
  const int N = a.lat.N, N2 = a.lat.N2;
  const int chain = blockIdx.x;
  const size_t off = (size_t)chain * N * N;
  const double *G00 = a.G00 + off, *Gtt = a.Gtt + off, *Gt0 = a.Gt0 + off, *G0t = a.G0t + off;
  const int8_t *sg = a.sig + (size_t)chain * a.T * N2;
  const int8_t *s_nt = sg + (size_t)(a.nt - 1) * N2, *s_1 = sg;
  const double RT = a.cc[chain].rt;
  const int d = threadIdx.x;
  if (d >= N) return;
  const int X = a.lat.xy[2 * d], Y = a.lat.xy[2 * d + 1];
  double *acc = a.acc + (size_t)chain * TAU_NOBS * a.T * N + (size_t)(a.nt - 1) * N + d;
  const size_t ks = (size_t)a.T * N;
  double spin = acc[TAU_SPIN * ks], ps = acc[TAU_PAIR_S * ks], pd = acc[TAU_PAIR_D * ks];
  double cs = acc[TAU_CHARGE_S * ks], cd = acc[TAU_CHARGE_D * ks], sgl = acc[TAU_SINGLE * ks];
  double cur = acc[TAU_CURXX * ks];
  for (int j = 0; j < N; ++j) {
    const int xj = a.lat.xy[2 * j], yj = a.lat.xy[2 * j + 1];
    const int xi = (xj - 1 + X) % a.Lx + 1, yi = (yj - 1 + Y) % a.Ly + 1;
    const int i = (yi - 1) * a.Lx + (xi - 1);
    spin = spin + (double)((int)s_nt[i] * (int)s_1[j]);
    const double g0t = G0t[j + (size_t)i * N];
    const double dg = g0t * g0t;
    pd = pd + (dg + dg) * 4.0;
    ps = ps + (dg + dg) * 4.0;
    const double x = 1.0 - Gtt[i + (size_t)i * N], y = 1.0 - G00[j + (size_t)j * N];
    const double tt = x + x, t0 = y + y;
    const double aa = g0t * Gt0[i + (size_t)j * N];
    cd = cd - aa - aa - aa - aa + (tt - tt) * (t0 - t0);
    cs = cs - aa - aa - aa - aa + (tt + tt) * (t0 + t0);
    const double term = Gt0[i + (size_t)j * N] + Gt0[i + (size_t)j * N];
    sgl = sgl + term + term;
    const int ip = a.lat.bond_p[2 * (2 * i) + 1], jp = a.lat.bond_p[2 * (2 * j) + 1];  // +x neighbours
    const double tc = -(G0t[j + (size_t)ip * N] * Gt0[i + (size_t)jp * N]) -
                      G0t[jp + (size_t)i * N] * Gt0[ip + (size_t)j * N] +
                      G0t[j + (size_t)i * N] * Gt0[ip + (size_t)jp * N] +
                      G0t[jp + (size_t)ip * N] * Gt0[i + (size_t)j * N];
    cur = cur - RT * RT * tc - RT * RT * tc;
    cur = cur - RT * RT * tc - RT * RT * tc;
  }
  acc[TAU_SPIN * ks] = spin;
  acc[TAU_PAIR_S * ks] = ps;
  acc[TAU_PAIR_D * ks] = pd;
  acc[TAU_CHARGE_S * ks] = cs;
  acc[TAU_CHARGE_D * ks] = cd;
  acc[TAU_SINGLE * ks] = sgl;
  acc[TAU_CURXX * ks] = cur;
}

// mat_big = [[invVRVL, diag(DUL)], [-diag(DUR), invULUR]]  (src/stabilization.f90:214-224)
__global__ void __launch_bounds__(256) k_assemble_big(const double *invVRVL, const double *invULUR,
                                                      const double *DUL, const double *DUR, double *big,
                                                      int N) {
  const int chain = blockIdx.y;
  const int N2 = 2 * N;
  const double *a = invVRVL + (size_t)chain * N * N, *b = invULUR + (size_t)chain * N * N;
  const double *dl = DUL + (size_t)chain * N, *dr = DUR + (size_t)chain * N;
  double *m = big + (size_t)chain * N2 * N2;
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < N2 * N2; idx += gridDim.x * blockDim.x) {
    const int c = idx / N2, r = idx - c * N2;
    double v = 0.0;
    if (r < N && c < N) v = a[r + (size_t)c * N];
    else if (r >= N && c >= N) v = b[(r - N) + (size_t)(c - N) * N];
    else if (r < N && c == r + N) v = dl[r];
    else if (r >= N && c == r - N) v = -dr[c];
    m[idx] = v;
  }
}

// n x n identity per chain
__global__ void __launch_bounds__(256) k_eye(double *X, int n) {
  const int chain = blockIdx.y;
  double *x = X + (size_t)chain * n * n;
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < n * n; idx += gridDim.x * blockDim.x)
    x[idx] = (idx / n == idx % n) ? 1.0 : 0.0;
}

// X (n x n) columns divided by d
__global__ void __launch_bounds__(256) k_div_cols(double *X, const double *d, int n) {
  const int chain = blockIdx.y;
  double *x = X + (size_t)chain * n * n;
  const double *dd = d + (size_t)chain * n;
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < n * n; idx += gridDim.x * blockDim.x)
    x[idx] = x[idx] / dd[idx / n];
}

// Wrap_tau tail (src/stabilization.f90:496-512): drift of Gr0t, Grt0, Grtt
// against the blocks of the recomputed 2N x 2N Green's function, then replace
// all four.  One workgroup per chain; drift[chain*6 + 0..2] max, [3..5] sums.
__global__ void __launch_bounds__(256) k_tau_replace(const double *tot, double *G00, double *G0t, double *Gt0,
                                                     double *Gtt, double *drift, int N) {
  const int chain = blockIdx.x;
  const int N2 = 2 * N;
  const double *t = tot + (size_t)chain * N2 * N2;
  const size_t off = (size_t)chain * N * N;
  double m0 = 0.0, m1 = 0.0, m2 = 0.0;
  for (int idx = threadIdx.x; idx < N * N; idx += blockDim.x) {
    const int c = idx / N, r = idx - c * N;
    const double n00 = t[r + (size_t)c * N2], n0t = t[r + (size_t)(c + N) * N2];
    const double nt0 = t[(r + N) + (size_t)c * N2], ntt = t[(r + N) + (size_t)(c + N) * N2];
    m0 = fmax(m0, fabs(n0t - G0t[off + idx]));
    m1 = fmax(m1, fabs(nt0 - Gt0[off + idx]));
    m2 = fmax(m2, fabs(ntt - Gtt[off + idx]));
    G00[off + idx] = n00;
    G0t[off + idx] = n0t;
    Gt0[off + idx] = nt0;
    Gtt[off + idx] = ntt;
  }
  for (int o = 32; o > 0; o >>= 1) {
    m0 = fmax(m0, __shfl_xor(m0, o, 64));
    m1 = fmax(m1, __shfl_xor(m1, o, 64));
    m2 = fmax(m2, __shfl_xor(m2, o, 64));
  }
  __shared__ double red[3][4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wid] = m0;
    red[1][wid] = m1;
    red[2][wid] = m2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double *dr = drift + (size_t)chain * 6;
    for (int q = 0; q < 3; ++q) {
      double mm = 0.0;
      for (int w = 0; w < 4; ++w) mm = fmax(mm, red[q][w]);
      if (mm > dr[q]) dr[q] = mm;
      dr[3 + q] += mm;
    }
  }
}

// acc *= z (Obs_tau_ave, src/obser_tau.f90:37-47)
__global__ void __launch_bounds__(256) k_scale(double *x, size_t n, double z) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    x[i] = z * x[i];
}
