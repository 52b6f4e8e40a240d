// Local link-flip sweep of one time slice (reference LocalK_metro,
// src/localK.f90:45-173, in the order of LocalK_prop_L/R, :347-398).
//
// One workgroup per chain.  Every proposal draws exactly one ranf, so the
// random number of proposal k of the slice is seed*a^(k+1) mod 2^31 (jump-ahead
// of ran_imada.f:55-68) and all 2N draws are known up front.  The proposals are
// evaluated speculatively in parallel against the current state; the first
// accepted one (in proposal order) is applied, its rank-2 update is recorded
// and folded into the cached 2x2 blocks of every bond, and evaluation resumes
// after it.  Proposals before the first accepted one were evaluated against the
// exact state the serial loop would have seen, so the accept/reject history is
// the serial one.  The recorded updates are applied to G in acceptance order:
// by this kernel whenever `kdelay` of them are pending (the delay depth), and
// the remaining ones by the next kinetic kernel (flush prologue).
#pragma once
#include "z2_common.hpp"

template <typename R>
struct FlipArgs {
  R *G;                   // chain stride N*N
  int8_t *sig;            // chain stride T*2N
  int32_t *seed;          // per chain
  R *Tbuf, *Vbuf;         // per chain stride upd_stride
  int *nacc;              // out: accepted updates in this slice, per chain
  unsigned long long *acc_count;  // per chain: [0] accepted, [1] proposed (cumulative)
  const ChainConst *cc;
  const uint32_t *apow;   // a^k mod 2^31, k = 0..2N
  long upd_stride;
  int T, nt;
  int descending;         // 1: bonds 2N..1 (left pass), 0: 1..2N (right pass)
  int mu_on_read;         // 1: G entries are read as (g*ep)*em (left pass, propT_L before flips)
  int kdelay;             // delay depth: pending updates before an in-kernel flush
  int *mu_done;           // out: 1 when an in-kernel flush applied the mu factors to G
  const uint32_t *force;  // optional: decisions to replay, bit k = proposal k, chain stride dwords
  uint32_t *record;       // optional: decisions taken, same layout
  int dwords;
  LatticeDev lat;
};

template <typename R>
__device__ __forceinline__ R flip_read(const R *G, int i, int j, int N, int mu, const ChainConst &c) {
  R x = G[i + (size_t)j * N];
  if (mu) {
    x = x * (R)c.ep;
    x = x * (R)c.em;
  }
  return x;
}

template <typename R>
__global__ void __launch_bounds__(512) k_flip(FlipArgs<R> a) {
  const int N = a.lat.N, N2 = a.lat.N2;
  const int chain = blockIdx.x;
  const ChainConst c = a.cc[chain];
  R *G = a.G + (size_t)chain * N * N;
  int8_t *sigc = a.sig + (size_t)chain * a.T * N2;
  int8_t *sig_t = sigc + (size_t)(a.nt - 1) * N2;
  const int tp = (a.nt - 2 + a.T) % a.T, tn = a.nt % a.T;  // 0-based neighbour slices
  const int8_t *sig_p = sigc + (size_t)tp * N2, *sig_n = sigc + (size_t)tn * N2;
  R *Tb = a.Tbuf + chain * a.upd_stride;
  R *Vb = a.Vbuf + chain * a.upd_stride;

  extern __shared__ unsigned char smem_raw[];
  double *dg = reinterpret_cast<double *>(smem_raw);  // N   : G_ii
  double *o12 = dg + N;                               // 2N  : G_p1p2 per bond
  double *o21 = o12 + N2;                             // 2N  : G_p2p1 per bond
  double *rnd = o21 + N2;                             // 2N  : random number per proposal index
  double *PT = rnd + N2;                              // kdelay x 4N : pending T0,T1,V0,V1
  int *sg = reinterpret_cast<int *>(PT + (size_t)4 * N * a.kdelay);  // 2N : current slice field
  __shared__ int s_kmin;
  __shared__ double s_pi[4], s_d[4];
  __shared__ int s_p[2];
  __shared__ int s_nacc, s_tot;
  __shared__ uint32_t s_dec[64];  // decision bits of this slice (2N <= 2048)

  int mu = a.mu_on_read;
  int mu_done = 0;
  for (int i = threadIdx.x; i < N; i += blockDim.x) dg[i] = (double)flip_read<R>(G, i, i, N, mu, c);
  const uint32_t seed0 = (uint32_t)a.seed[chain] & 0x7fffffffu;
  for (int b = threadIdx.x; b < N2; b += blockDim.x) {
    int p1 = a.lat.bond_p[2 * b], p2 = a.lat.bond_p[2 * b + 1];
    o12[b] = (double)flip_read<R>(G, p1, p2, N, mu, c);
    o21[b] = (double)flip_read<R>(G, p2, p1, N, mu, c);
    sg[b] = (int)sig_t[b];
    // proposal index k = b: draw number k+1 of this slice
    uint64_t x = (uint64_t)seed0 * (uint64_t)a.apow[b + 1];
    rnd[b] = (double)(uint32_t)(x & 0x7fffffffull) / 2147483648.0;
  }
  if (threadIdx.x == 0) {
    s_nacc = 0;
    s_tot = 0;
  }
  for (int q = threadIdx.x; q < 64; q += blockDim.x) s_dec[q] = 0u;
  const uint32_t *frc = a.force ? a.force + (size_t)chain * a.dwords : nullptr;
  __syncthreads();

  int start = 0;
  while (start < N2) {
    if (threadIdx.x == 0) s_kmin = N2;
    __syncthreads();
    for (int k = start + threadIdx.x; k < N2; k += blockDim.x) {
      const int b = a.descending ? (N2 - 1 - k) : k;
      const int p1 = a.lat.bond_p[2 * b], p2 = a.lat.bond_p[2 * b + 1];
      const int s_old = sg[b];
      const double *D = c.delta[s_old > 0 ? 0 : 1];
      const double A11 = 1.0 - dg[p1], A21 = 0.0 - o21[b];
      const double A12 = 0.0 - o12[b], A22 = 1.0 - dg[p2];
      const double P11 = 1.0 + (D[0] * A11 + D[2] * A21);
      const double P21 = 0.0 + (D[1] * A11 + D[3] * A21);
      const double P12 = 0.0 + (D[0] * A12 + D[2] * A22);
      const double P22 = 1.0 + (D[1] * A12 + D[3] * A22);
      const double det = P11 * P22 - P12 * P21;
      // boson ratio: plaquette sums (src/fields.f90:66-101) and temporal term (:104-117)
      const int *bb = a.lat.bond_bonds + 6 * b;
      const int s_new = -s_old;
      int tn1 = s_new * sg[bb[0]] * sg[bb[1]] * sg[bb[2]];
      int tn2 = s_new * sg[bb[3]] * sg[bb[4]] * sg[bb[5]];
      int dS = 2 * (tn1 + tn2);  // tmp_new - tmp_old with tmp_old = -tmp_new
      int nb = (int)sig_p[b] + (int)sig_n[b];
      int dT = 2 * s_new * nb;
      const double rb = c.boson[(dS / 4 + 1) * 3 + (dT / 4 + 1)];
      const double ratio = (det * det) * rb;
      const bool acc = frc ? ((frc[k >> 5] >> (k & 31)) & 1u) != 0u : fabs(ratio) > rnd[k];
      if (acc) {
        atomicMin(&s_kmin, k);
        break;  // later k of this thread are larger
      }
    }
    __syncthreads();
    const int kmin = s_kmin;
    if (kmin >= N2) break;
    const int b = a.descending ? (N2 - 1 - kmin) : kmin;
    const int na = s_nacc;
    if (threadIdx.x == 0) {
      const int p1 = a.lat.bond_p[2 * b], p2 = a.lat.bond_p[2 * b + 1];
      const int s_old = sg[b];
      const double *D = c.delta[s_old > 0 ? 0 : 1];
      const double A11 = 1.0 - dg[p1], A21 = 0.0 - o21[b];
      const double A12 = 0.0 - o12[b], A22 = 1.0 - dg[p2];
      const double P11 = 1.0 + (D[0] * A11 + D[2] * A21);
      const double P21 = 0.0 + (D[1] * A11 + D[3] * A21);
      const double P12 = 0.0 + (D[0] * A12 + D[2] * A22);
      const double P22 = 1.0 + (D[1] * A12 + D[3] * A22);
      const double det = P11 * P22 - P12 * P21;
      s_pi[0] = P22 / det;     // Pinv(1,1)
      s_pi[1] = (-P21) / det;  // Pinv(2,1)
      s_pi[2] = (-P12) / det;  // Pinv(1,2)
      s_pi[3] = P11 / det;     // Pinv(2,2)
      for (int q = 0; q < 4; ++q) s_d[q] = D[q];
      s_p[0] = p1;
      s_p[1] = p2;
      sg[b] = -s_old;
      s_dec[kmin >> 5] |= 1u << (kmin & 31);
      s_tot = s_tot + 1;
    }
    __syncthreads();
    double *TA = PT + (size_t)4 * N * na;  // slot of this accept: T0, T1, V0, V1
    {
      const int p1 = s_p[0], p2 = s_p[1];
      const double Pi11 = s_pi[0], Pi21 = s_pi[1], Pi12 = s_pi[2], Pi22 = s_pi[3];
      const double D11 = s_d[0], D21 = s_d[1], D12 = s_d[2], D22 = s_d[3];
      for (int i = threadIdx.x; i < N; i += blockDim.x) {
        // current G(i,p1), G(i,p2), G(p1,i), G(p2,i): pending accepts applied in order
        double c1 = (double)flip_read<R>(G, i, p1, N, mu, c);
        double c2 = (double)flip_read<R>(G, i, p2, N, mu, c);
        double r1 = (double)flip_read<R>(G, p1, i, N, mu, c);
        double r2 = (double)flip_read<R>(G, p2, i, N, mu, c);
        for (int k = 0; k < na; ++k) {
          const double *P = PT + (size_t)4 * N * k;
          double ti0 = P[i], ti1 = P[N + i];
          double tp10 = P[p1], tp11 = P[N + p1], tp20 = P[p2], tp21 = P[N + p2];
          double vi0 = P[2 * N + i], vi1 = P[3 * N + i];
          double vp10 = P[2 * N + p1], vp11 = P[3 * N + p1], vp20 = P[2 * N + p2], vp21 = P[3 * N + p2];
          c1 = c1 - (ti0 * vp10 + ti1 * vp11);
          c2 = c2 - (ti0 * vp20 + ti1 * vp21);
          r1 = r1 - (tp10 * vi0 + tp11 * vi1);
          r2 = r2 - (tp20 * vi0 + tp21 * vi1);
        }
        double t0 = c1 * Pi11 + c2 * Pi21;
        double t1 = c1 * Pi12 + c2 * Pi22;
        double v0 = -(D11 * r1) - D12 * r2;
        double v1 = -(D21 * r1) - D22 * r2;
        if (i == p1) {
          v0 = v0 + D11;
          v1 = v1 + D21;
        }
        if (i == p2) {
          v0 = v0 + D12;
          v1 = v1 + D22;
        }
        TA[i] = t0;
        TA[N + i] = t1;
        TA[2 * N + i] = v0;
        TA[3 * N + i] = v1;
        Tb[(2 * na) * N + i] = (R)t0;
        Tb[(2 * na + 1) * N + i] = (R)t1;
        Vb[(2 * na) * N + i] = (R)v0;
        Vb[(2 * na + 1) * N + i] = (R)v1;
      }
    }
    __syncthreads();
    // fold the update into the cached 2x2 blocks
    const double *VA = TA + 2 * N;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
      double d = TA[i] * VA[i] + TA[N + i] * VA[N + i];
      dg[i] = dg[i] - d;
    }
    for (int bb2 = threadIdx.x; bb2 < N2; bb2 += blockDim.x) {
      int q1 = a.lat.bond_p[2 * bb2], q2 = a.lat.bond_p[2 * bb2 + 1];
      double d12 = TA[q1] * VA[q2] + TA[N + q1] * VA[N + q2];
      double d21 = TA[q2] * VA[q1] + TA[N + q2] * VA[N + q1];
      o12[bb2] = o12[bb2] - d12;
      o21[bb2] = o21[bb2] - d21;
    }
    if (threadIdx.x == 0) s_nacc = na + 1;
    __syncthreads();
    if (na + 1 >= a.kdelay) {
      // flush the pending updates into G in acceptance order (and mu, once)
      const int np = na + 1;
      for (int idx = threadIdx.x; idx < N * N; idx += blockDim.x) {
        const int col = idx / N, r = idx - col * N;
        R x = G[idx];
        if (mu) {
          x = x * (R)c.ep;
          x = x * (R)c.em;
        }
        for (int k = 0; k < np; ++k) {
          const double *P = PT + (size_t)4 * N * k;
          R t0 = (R)P[r], t1 = (R)P[N + r];
          R v0 = (R)P[2 * N + col], v1 = (R)P[3 * N + col];
          R pp = t0 * v0;
          R qq = t1 * v1;
          R dd = pp + qq;
          x = x - dd;
        }
        G[idx] = x;
      }
      if (mu) mu_done = 1;
      mu = 0;
      __syncthreads();
      if (threadIdx.x == 0) s_nacc = 0;
      __syncthreads();
    }
    start = kmin + 1;
  }
  __syncthreads();
  for (int b = threadIdx.x; b < N2; b += blockDim.x) sig_t[b] = (int8_t)sg[b];
  if (a.record)
    for (int q = threadIdx.x; q < a.dwords; q += blockDim.x) a.record[(size_t)chain * a.dwords + q] = s_dec[q];
  if (threadIdx.x == 0) {
    a.nacc[chain] = s_nacc;
    a.mu_done[chain] = mu_done;
    uint64_t x = (uint64_t)seed0 * (uint64_t)a.apow[N2];
    a.seed[chain] = (int32_t)(uint32_t)(x & 0x7fffffffull);
    a.acc_count[2 * chain] += (unsigned long long)s_tot;
    a.acc_count[2 * chain + 1] += (unsigned long long)N2;
  }
}
