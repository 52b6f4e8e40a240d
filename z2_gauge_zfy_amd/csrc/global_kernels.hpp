// Wolff global update of the temporal line of one bond (reference
// src/global_update.f90:65-337, src/globalK.f90:15-151), the boson-only warm-up
// sweep (LocalK_therm, src/localK.f90:400-420), and the per-chain block copies
// used to back up, restore and reorder chain state.
//
// The cluster and the accept/reject step are serial per chain and draw from the
// chain's own ranf stream, so one thread serves one chain.  The fermion ratio
// of a trial is accumulated by the flip kernel run in forced mode (the cluster
// bond is "accepted" at the cluster's slices, every other proposal rejected)
// inside an ordinary propagation pass.
#pragma once
#include "z2_common.hpp"

#define Z2_RATIO_EPS 1.0e-300  // src/globalK.f90:9

// ranf, lib/package_of_matrix/Ran/ran_imada.f:55-68: iq <- iq * 5^11 mod 2^31
__device__ __forceinline__ double dev_ranf(uint32_t &iq) {
  iq = (uint32_t)(((uint64_t)iq * 48828125ull) & 0x7fffffffull);
  return (double)iq / 2147483648.0;
}

// nranf, src/calc_basic.f90:113-121: nint(ranf * n + 0.5) clamped to [1, n]
__device__ __forceinline__ int dev_nranf(uint32_t &iq, int n) {
  const double x = dev_ranf(iq) * (double)n + 0.5;
  const int r = (int)round(x);  // half away from zero, as Fortran nint
  return r < 1 ? 1 : (r > n ? n : r);
}

struct GlobArgs {
  int8_t *sig;                // chain stride T * 2N
  int32_t *seed;              // per chain
  const ChainConst *cc;       // per chain
  LatticeDev lat;
  int T, nch;
  int *bond;                  // per chain: bond of the trial (0-based)
  int *mask;                  // per chain T: 1 at the slices of the cluster
  int *queue;                 // per chain T: breadth-first queue of slices (1-based)
  double *lr;                 // per chain [2]: fermion log ratio, boson log ratio of the trial
  int *acc;                   // per chain: 1 when the last trial was accepted
  unsigned long long *cnt;    // per chain [3]: accepted, proposed, sum of cluster sizes
  double *sz;                 // per chain [2]: cluster sizes of this global sweep, wolff_size_cluster
};

// Wolff_flip + build_time_cluster + compute_space_log_ratio
// (src/global_update.f90:65-172): picks a bond and a seed slice, grows the
// cluster along imaginary time through equal neighbouring links with
// probability p_time, and sums the plaquette log ratio over the cluster.
__global__ void __launch_bounds__(64) k_wolff_cluster(GlobArgs a) {
  const int chain = blockIdx.x * blockDim.x + threadIdx.x;
  if (chain >= a.nch) return;
  const int T = a.T, N2 = a.lat.N2;
  const ChainConst *cp = a.cc + chain;
  const int8_t *s = a.sig + (size_t)chain * T * N2;
  int *mask = a.mask + (size_t)chain * T, *queue = a.queue + (size_t)chain * T;
  uint32_t iq = (uint32_t)a.seed[chain];
  const int bond = dev_nranf(iq, N2);  // 1-based
  const double p_time = cp->ptime;
  for (int t = 0; t < T; ++t) {
    mask[t] = 0;
    queue[t] = 0;
  }
  const int q0 = dev_nranf(iq, T);
  queue[0] = q0;
  mask[q0 - 1] = 1;
  const int8_t *sb = s + (bond - 1);  // sigma(bond, t) at sb[(t-1) * N2]
  int head = 0, tail = 0;
  while (head <= tail) {
    const int cur = queue[head++];
    for (int off = -1; off <= 1; off += 2) {
      int nb = cur + off;
      nb = nb < 1 ? nb + T : (nb > T ? nb - T : nb);
      if (mask[nb - 1]) continue;
      if (sb[(size_t)(nb - 1) * N2] == sb[(size_t)(cur - 1) * N2]) {
        if (p_time > 0.0 && dev_ranf(iq) < p_time) {
          queue[++tail] = nb;  // every queued slice was unmasked: tail < T
          mask[nb - 1] = 1;
        }
      }
    }
  }
  int size = 0;
  for (int t = 0; t < T; ++t) size += mask[t];
  // plaquette_log_ratio over the cluster, before the flip (src/global_update.f90:135-172)
  double lr = 0.0;
  const double alpha = cp->walpha;
  if (alpha != 0.0) {
    const int *bb = a.lat.bond_bonds + 6 * (bond - 1);
    for (int t = 1; t <= T; ++t) {
      if (!mask[t - 1]) continue;
      const int8_t *st = s + (size_t)(t - 1) * N2;
      for (int h = 0; h < 2; ++h) {
        double prod = (double)st[bond - 1];
        for (int k = 0; k < 3; ++k) prod = prod * (double)st[bb[3 * h + k]];
        lr = lr + -2.0 * alpha * prod;
      }
    }
  }
  a.bond[chain] = bond - 1;
  a.lr[2 * chain] = 0.0;
  a.lr[2 * chain + 1] = lr;
  a.seed[chain] = (int32_t)iq;
  a.cnt[3 * chain + 2] += (unsigned long long)size;
  a.sz[2 * chain] = a.sz[2 * chain] + (double)size;
}

// the Metropolis test of Global_sweep_L/R (src/global_update.f90:229-311)
__global__ void __launch_bounds__(64) k_wolff_accept(GlobArgs a) {
  const int chain = blockIdx.x * blockDim.x + threadIdx.x;
  if (chain >= a.nch) return;
  uint32_t iq = (uint32_t)a.seed[chain];
  const double total = a.lr[2 * chain] + a.lr[2 * chain + 1];
  const double r = dev_ranf(iq);
  const double lhs = total < 0.0 ? total : 0.0;
  const double rr = r > Z2_RATIO_EPS ? r : Z2_RATIO_EPS;
  const int acc = lhs >= log(rr) ? 1 : 0;
  a.cnt[3 * chain + 1] += 1ull;
  a.cnt[3 * chain] += (unsigned long long)acc;
  a.acc[chain] = acc;
  a.seed[chain] = (int32_t)iq;
}

// end of Global_sweep: wolff_size_cluster += (sum of cluster sizes) / Nglobal
__global__ void __launch_bounds__(64) k_wolff_finish(GlobArgs a, int nglobal) {
  const int chain = blockIdx.x * blockDim.x + threadIdx.x;
  if (chain >= a.nch) return;
  const double sc = a.sz[2 * chain] / (double)nglobal;
  a.sz[2 * chain + 1] = a.sz[2 * chain + 1] + sc;
  a.sz[2 * chain] = 0.0;
}

// LocalK_therm over every (nt, ii), nt outer (src/localK.f90:400-420, in the
// loop of Local_sweep_therm, src/local_sweep.f90:73-85).  One workgroup per chain: the slices nt-1,
// nt, nt+1 are staged in LDS and one thread runs the serial proposals of slice nt.
struct ThermArgs {
  int8_t *sig;
  int32_t *seed;
  const ChainConst *cc;
  LatticeDev lat;
  int T;
  unsigned long long *cnt;  // per chain [2]: accepted, proposed
};

__global__ void __launch_bounds__(256) k_therm(ThermArgs a) {
  const int chain = blockIdx.x;
  const int T = a.T, N2 = a.lat.N2;
  const ChainConst *cp = a.cc + chain;
  int8_t *s = a.sig + (size_t)chain * T * N2;
  extern __shared__ unsigned char smem_raw[];
  int8_t *buf = reinterpret_cast<int8_t *>(smem_raw);  // 3 x 2N
  double *bos = reinterpret_cast<double *>(smem_raw + ((3 * (size_t)N2 + 15) & ~(size_t)15));
  if (threadIdx.x < 9) bos[threadIdx.x] = cp->boson[threadIdx.x];
  uint32_t iq = (uint32_t)a.seed[chain];
  unsigned long long nacc = 0;
  for (int nt = 1; nt <= T; ++nt) {
    const int tp = nt == 1 ? T : nt - 1, tn = nt == T ? 1 : nt + 1;
    // LDS slots: 1 = slice nt; the neighbours alias it (T = 1) or each other (T = 2)
    const int sp = tp == nt ? 1 : 0;
    const int sn = tn == nt ? 1 : (tn == tp ? sp : 2);
    __syncthreads();
    for (int b = threadIdx.x; b < N2; b += blockDim.x) {
      buf[N2 + b] = s[(size_t)(nt - 1) * N2 + b];
      if (sp == 0) buf[b] = s[(size_t)(tp - 1) * N2 + b];
      if (sn == 2) buf[2 * N2 + b] = s[(size_t)(tn - 1) * N2 + b];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int8_t *sg = buf + N2;
      const int8_t *gp = buf + (size_t)sp * N2, *gn = buf + (size_t)sn * N2;
      for (int b = 0; b < N2; ++b) {
        // m_bosonratio_local (src/fields.f90:40-119) from the 3 x 3 table, as in k_flip
        const int *bb = a.lat.bond_bonds + 6 * b;
        const int s_old = sg[b], s_new = -s_old;
        const int tn1 = s_new * (int)sg[bb[0]] * (int)sg[bb[1]] * (int)sg[bb[2]];
        const int tn2 = s_new * (int)sg[bb[3]] * (int)sg[bb[4]] * (int)sg[bb[5]];
        const int dS = 2 * (tn1 + tn2);
        const int dT = 2 * s_new * ((int)gp[b] + (int)gn[b]);
        const double rb = bos[(dS / 4 + 1) * 3 + (dT / 4 + 1)];
        const double r = dev_ranf(iq);
        if (rb > r) {
          sg[b] = (int8_t)s_new;
          ++nacc;
        }
      }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < N2; b += blockDim.x) s[(size_t)(nt - 1) * N2 + b] = buf[N2 + b];
  }
  if (threadIdx.x == 0) {
    a.seed[chain] = (int32_t)iq;
    a.cnt[2 * chain] += nacc;
    a.cnt[2 * chain + 1] += (unsigned long long)T * N2;
  }
}

// dst chain j <- src chain j (or src chain perm[j]) for chains j of a range, in
// every slot of a slot-major array; chains whose skip flag is set are left alone.
// W: the copy word (16, 8 or 4 bytes, dividing the per-chain size).
template <typename W>
__global__ void __launch_bounds__(256) k_copy_chains(const W *src, W *dst, size_t words, size_t slot_words,
                                                     const int *perm, const int *skip) {
  const int j = blockIdx.y, s = blockIdx.z;
  if (skip && skip[j]) return;
  const int js = perm ? perm[j] : j;
  const W *a = src + (size_t)s * slot_words + (size_t)js * words;
  W *b = dst + (size_t)s * slot_words + (size_t)j * words;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (size_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}
