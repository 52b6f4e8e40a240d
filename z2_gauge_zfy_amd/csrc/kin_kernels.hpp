// Checkerboard kinetic propagator B = E1 E2 E3 E4 E4 E3 E2 E1 applied from the
// left (row-pair rotations, reference src/operator_kin.f90:61-120) or from the
// right (column-pair rotations, src/operator_kin.f90:122-179), all 8 stages
// fused in one pass through LDS.  Optional prologue: scalar mu factors
// (src/non_interact.f90:34-82) and the delayed rank-2 updates of the flip
// kernel, applied one after the other in acceptance order -- the same
// arithmetic as the reference's immediate updates (src/localK.f90:150-152).
// Optional epilogue: one scalar factor.
#pragma once
#include "z2_common.hpp"

// scalar factor selectors
enum : int { FAC_NONE = 0, FAC_EP = 1, FAC_EM = 2 };

template <typename R>
struct KinMat {
  R *mat;         // chain stride N*N
  int nflag;      // +1 / -1 rotation, 0: none
  int flush;      // apply delayed updates
  int pre1, pre2; // factors before the flush
  int pre_cond;   // skip pre1/pre2 for chains whose flip kernel already applied them
  int mid1, mid2; // factors after the flush
  int post, post2; // factors after the rotation
};

template <typename R>
struct KinArgs {
  KinMat<R> m[4];        // grid.z selects the matrix
  const int8_t *sig;     // sigma base: chain stride T*2N, slice (nt-1)*2N
  int T, nt;
  const ChainConst *cc;  // per chain
  const R *Tbuf;         // delayed updates: chain stride upd_stride, [(k*2+c)*N + row]
  const R *Vbuf;         //                                          [(k*2+c)*N + col]
  const int *nacc;       // accepted updates per chain
  const int *mu_done;    // per chain: mu factors already applied by the flip kernel
  long upd_stride;
  LatticeDev lat;
};

// Arithmetic: the matrices are stored in R (FP64, or FP32 in the mixed mode) and
// transformed in FP64.  A stored element is rounded to R once per procedure of
// the reference: the flush of the pending updates and the 8 rotation stages run
// in FP64 on the loaded value; each scalar mu factor (propT_*, a procedure of its
// own) is applied as the product in R, so a fused launch rounds exactly where the
// reference's procedure sequence does.  For R = double nothing is rounded beyond
// the FP64 operations themselves.
template <typename R>
__device__ __forceinline__ double to_r(double x) {
  return (double)(R)x;
}

// x times the scalar factor sel, as the product in R (exact in FP64 before the
// rounding: both operands have R's precision)
template <typename R>
__device__ __forceinline__ double apply_fac(double x, int sel, const ChainConst &c) {
  if (sel == FAC_EP) return to_r<R>(x * (double)(R)c.ep);
  if (sel == FAC_EM) return to_r<R>(x * (double)(R)c.em);
  return x;
}

// Prologue applied to element (r, col) of a chain's matrix: scalar factors, then
// the pending updates in acceptance order (FP64), then scalar factors on the
// rounded result.
template <typename R>
__device__ __forceinline__ double kin_prologue(R xin, int r, int col, const KinMat<R> &km,
                                               const KinArgs<R> &a, const ChainConst &c, int chain, bool pre_on) {
  double x = (double)xin;
  if (pre_on) {
    x = apply_fac<R>(x, km.pre1, c);
    x = apply_fac<R>(x, km.pre2, c);
  }
  if (km.flush) {
    const int N = a.lat.N;
    const int na = a.nacc[chain];
    const R *T = a.Tbuf + chain * a.upd_stride;
    const R *V = a.Vbuf + chain * a.upd_stride;
    for (int k = 0; k < na; ++k) {
      double t0 = (double)T[(2 * k) * N + r], t1 = (double)T[(2 * k + 1) * N + r];
      double v0 = (double)V[(2 * k) * N + col], v1 = (double)V[(2 * k + 1) * N + col];
      double p = t0 * v0;
      double q = t1 * v1;
      double d = p + q;
      x = x - d;
    }
  }
  if (km.mid1 || km.mid2) {
    x = to_r<R>(x);
    x = apply_fac<R>(x, km.mid1, c);
    x = apply_fac<R>(x, km.mid2, c);
  }
  return x;
}

// the stored value of a transformed element with the epilogue factors
template <typename R>
__device__ __forceinline__ R kin_epilogue(double x, const KinMat<R> &km, const ChainConst &c) {
  if (km.post || km.post2) {
    x = to_r<R>(x);
    x = apply_fac<R>(x, km.post, c);
    x = apply_fac<R>(x, km.post2, c);
  }
  return (R)x;
}

// threads per block of every kinetic launch (the kernels' launch bounds)
constexpr int KIN_NT = 256;

// Rotation table of one slice in LDS: for each of the 4 checkerboard groups and
// each of its N/2 bonds, the two sites and the sign of nflag * sigma packed in
// one int (sites < 2048), so a rotation needs one LDS read instead of a chain of
// dependent global reads.  Entry e = g N/2 + k (bond k of group g); the block's
// nt threads build it two entries per thread per step, both entries' dependent
// reads in flight together (the second index clamped to the last entry: past
// the end it rewrites that entry with its own value).
__device__ __forceinline__ void kin_tab_build(int *tab, const LatticeDev &lat, const int8_t *s, int nflag, int n,
                                              int tid, int nt) {
  for (int e0 = tid; e0 < n; e0 += 2 * nt) {
    const int e1 = min(e0 + nt, n - 1);
    const int b0 = lat.grp[e0], b1 = lat.grp[e1];
    const int p10 = lat.bond_p[2 * b0], p20 = lat.bond_p[2 * b0 + 1], s0 = s[b0];
    const int p11 = lat.bond_p[2 * b1], p21 = lat.bond_p[2 * b1 + 1], s1 = s[b1];
    tab[e0] = p10 | (p20 << 11) | ((nflag * s0 > 0) ? (1 << 22) : 0);
    tab[e1] = p11 | (p21 << 11) | ((nflag * s1 > 0) ? (1 << 22) : 0);
  }
}

// Mat <- B^{nflag} Mat : each column is independent; a block owns CB columns.
// A thread owns one bond of a group and a set of the block's columns: its
// rotation data is read once per stage.
// grid: (ceil(N/CB), n_chains, n_mats)
// NC: the site count when fixed at compile time (0: a.lat.N)
template <typename R, int CB, int NC = 0>
__global__ void __launch_bounds__(256) k_kin_rows(KinArgs<R> a) {
  extern __shared__ unsigned char smem_raw[];
  double *tile = reinterpret_cast<double *>(smem_raw);         // [CB][N]
  int *tab = reinterpret_cast<int *>(tile + (size_t)CB * a.lat.N);  // [4][N/2]
  const int N = NC ? NC : a.lat.N, N2 = 2 * N, half = N / 2, nt = blockDim.x, tid = threadIdx.x;
  const int chain = blockIdx.y;
  const KinMat<R> km = a.m[blockIdx.z];
  R *M = km.mat + (size_t)chain * N * N;
  const int col0 = blockIdx.x * CB;
  const int ncol = min(CB, N - col0);
  const ChainConst c = a.cc[chain];
  if (km.nflag != 0) {
    const int8_t *s = a.sig + (size_t)chain * a.T * N2 + (size_t)(a.nt - 1) * N2;
    kin_tab_build(tab, a.lat, s, km.nflag, 4 * half, tid, nt);
  }
  // the tile, the reads of a thread all in flight before the prologue (sizes fixed
  // at compile time: the block has KIN_NT threads)
  const bool pre_on = !(km.pre_cond && a.mu_done[chain]);
  if constexpr (NC > 0) {
    constexpr int RPT = (NC + KIN_NT - 1) / KIN_NT;
    R raw[CB][RPT];
#pragma unroll
    for (int cc = 0; cc < CB; ++cc)
#pragma unroll
      for (int j = 0; j < RPT; ++j)
        raw[cc][j] = M[min(tid + j * KIN_NT, NC - 1) + (size_t)(col0 + min(cc, ncol - 1)) * NC];
#pragma unroll
    for (int cc = 0; cc < CB; ++cc)
#pragma unroll
      for (int j = 0; j < RPT; ++j) {
        const int r = tid + j * KIN_NT;
        if (cc < ncol && r < NC) tile[cc * NC + r] = kin_prologue<R>(raw[cc][j], r, col0 + cc, km, a, c, chain, pre_on);
      }
  } else {
    for (int cc = 0; cc < ncol; ++cc)
      for (int r = tid; r < N; r += nt)
        tile[cc * N + r] = kin_prologue<R>(M[r + (size_t)(col0 + cc) * N], r, col0 + cc, km, a, c, chain, pre_on);
  }
  __syncthreads();
  if (km.nflag != 0) {
    const double Cp = c.Cp, Cm = c.Cm, Sp = c.Sp, Sm = c.Sm;
    const int ngrp = half >= nt ? 1 : nt / half;  // column groups
    const int items = half * ngrp;
    for (int st = 0; st < 8; ++st) {
      const int *t = tab + kin_group_of_stage(st) * half;
      for (int it = tid; it < items; it += nt) {
        const int k = it % half, cg = it / half;
        const int e = t[k];
        const int p1 = e & 2047, p2 = (e >> 11) & 2047;
        const double C = (e >> 22) ? Cp : Cm, S = (e >> 22) ? Sp : Sm;
        for (int cc = cg; cc < ncol; cc += ngrp) {
          double x1 = tile[cc * N + p1], x2 = tile[cc * N + p2];
          double v1 = C * x1 + S * x2;
          double v2 = S * x1 + C * x2;
          tile[cc * N + p1] = v1;
          tile[cc * N + p2] = v2;
        }
      }
      __syncthreads();
    }
  }
  for (int cc = 0; cc < ncol; ++cc)
    for (int r = tid; r < N; r += nt) M[r + (size_t)(col0 + cc) * N] = kin_epilogue<R>(tile[cc * N + r], km, c);
}

// Mat <- Mat B^{nflag} : each row is independent; a block owns RB rows.  Thread
// tid works on row tid % RB of the block and on bonds tid / RB + j * (nt / RB).
template <typename R, int RB, int NC = 0>
__global__ void __launch_bounds__(256) k_kin_cols(KinArgs<R> a) {
  extern __shared__ unsigned char smem_raw[];
  double *tile = reinterpret_cast<double *>(smem_raw);         // [N][RB]
  int *tab = reinterpret_cast<int *>(tile + (size_t)RB * a.lat.N);  // [4][N/2]
  const int N = NC ? NC : a.lat.N, N2 = 2 * N, half = N / 2, nt = blockDim.x, tid = threadIdx.x;
  const int chain = blockIdx.y;
  const KinMat<R> km = a.m[blockIdx.z];
  R *M = km.mat + (size_t)chain * N * N;
  const int row0 = blockIdx.x * RB;
  const int nrow = min(RB, N - row0);
  const ChainConst c = a.cc[chain];
  const int r = tid % RB, cstep = nt / RB;
  if (km.nflag != 0) {
    const int8_t *s = a.sig + (size_t)chain * a.T * N2 + (size_t)(a.nt - 1) * N2;
    kin_tab_build(tab, a.lat, s, km.nflag, 4 * half, tid, nt);
  }
  // the tile, the reads of a thread all in flight before the prologue
  const bool pre_on = !(km.pre_cond && a.mu_done[chain]);
  if constexpr (NC > 0) {
    constexpr int CSTEP = KIN_NT / RB, CPT = (NC + CSTEP - 1) / CSTEP;
    R raw[CPT];
    const int rr = row0 + min(r, nrow - 1);
#pragma unroll
    for (int j = 0; j < CPT; ++j) raw[j] = M[rr + (size_t)min(tid / RB + j * CSTEP, NC - 1) * NC];
    if (r < nrow)
#pragma unroll
      for (int j = 0; j < CPT; ++j) {
        const int col = tid / RB + j * CSTEP;
        if (col < NC) tile[col * RB + r] = kin_prologue<R>(raw[j], row0 + r, col, km, a, c, chain, pre_on);
      }
  } else if (r < nrow) {
    for (int col = tid / RB; col < N; col += cstep)
      tile[col * RB + r] = kin_prologue<R>(M[(row0 + r) + (size_t)col * N], row0 + r, col, km, a, c, chain, pre_on);
  }
  __syncthreads();
  if (km.nflag != 0) {
    const double Cp = c.Cp, Cm = c.Cm, Sp = c.Sp, Sm = c.Sm;
    for (int st = 0; st < 8; ++st) {
      const int *t = tab + kin_group_of_stage(st) * half;
      if (r < nrow)
        for (int k = tid / RB; k < half; k += cstep) {
          const int e = t[k];
          const int p1 = e & 2047, p2 = (e >> 11) & 2047;
          const double C = (e >> 22) ? Cp : Cm, S = (e >> 22) ? Sp : Sm;
          double x1 = tile[p1 * RB + r], x2 = tile[p2 * RB + r];
          double u1 = x1 * C + x2 * S;
          double u2 = x1 * S + x2 * C;
          tile[p1 * RB + r] = u1;
          tile[p2 * RB + r] = u2;
        }
      __syncthreads();
    }
  }
  if (r < nrow)
    for (int col = tid / RB; col < N; col += cstep)
      M[(row0 + r) + (size_t)col * N] = kin_epilogue<R>(tile[col * RB + r], km, c);
}

// Elementwise prologue only (flush and/or mu factors) over whole matrices.
template <typename R>
__global__ void __launch_bounds__(256) k_kin_elem(KinArgs<R> a) {
  const int N = a.lat.N;
  const int chain = blockIdx.y;
  const KinMat<R> km = a.m[blockIdx.z];
  R *M = km.mat + (size_t)chain * N * N;
  const ChainConst c = a.cc[chain];
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < N * N;
       idx += gridDim.x * blockDim.x) {
    int col = idx / N, r = idx - col * N;
    M[idx] = kin_epilogue<R>(kin_prologue<R>(M[idx], r, col, km, a, c, chain, !(km.pre_cond && a.mu_done[chain])), km, c);
  }
}

// Several slices' wraps of a matrix nothing reads between them (the propagators
// UUL, UUR between two stabilizations): the slices nt0, nt0 + step, ... (ns of
// them) applied to one tile in LDS in that order, the tile read and written once.
// Every element sees the operations of ns launches of k_kin_cols / k_kin_rows
// with the same KinMat (scalar factors before and after the rotations, no flush),
// including the rounding to R between slices, so the result is theirs bit for bit.
template <typename R>
__device__ __forceinline__ double kin_multi_pre(double x, const KinMat<R> &km, const ChainConst &c) {
  x = apply_fac<R>(x, km.pre1, c);
  return apply_fac<R>(x, km.pre2, c);
}

template <typename R, int CB, int NC = 0>
__global__ void __launch_bounds__(256) k_kin_rows_multi(KinArgs<R> a, int nt0, int ns, int step) {
  extern __shared__ unsigned char smem_raw[];
  double *tile = reinterpret_cast<double *>(smem_raw);              // [CB][N]
  int *tab = reinterpret_cast<int *>(tile + (size_t)CB * a.lat.N);  // [4][N/2]
  const int N = NC ? NC : a.lat.N, N2 = 2 * N, half = N / 2, nt = blockDim.x, tid = threadIdx.x;
  const int chain = blockIdx.y;
  const KinMat<R> km = a.m[blockIdx.z];
  R *M = km.mat + (size_t)chain * N * N;
  const int col0 = blockIdx.x * CB;
  const int ncol = min(CB, N - col0);
  const ChainConst c = a.cc[chain];
  if constexpr (NC > 0) {  // the reads of a thread all in flight
    constexpr int RPT = (NC + KIN_NT - 1) / KIN_NT;
    R raw[CB][RPT];
#pragma unroll
    for (int cc = 0; cc < CB; ++cc)
#pragma unroll
      for (int j = 0; j < RPT; ++j)
        raw[cc][j] = M[min(tid + j * KIN_NT, NC - 1) + (size_t)(col0 + min(cc, ncol - 1)) * NC];
#pragma unroll
    for (int cc = 0; cc < CB; ++cc)
#pragma unroll
      for (int j = 0; j < RPT; ++j) {
        const int r = tid + j * KIN_NT;
        if (cc < ncol && r < NC) tile[cc * NC + r] = (double)raw[cc][j];
      }
  } else {
    for (int cc = 0; cc < ncol; ++cc)
      for (int r = tid; r < N; r += nt) tile[cc * N + r] = (double)M[r + (size_t)(col0 + cc) * N];
  }
  const double Cp = c.Cp, Cm = c.Cm, Sp = c.Sp, Sm = c.Sm;
  const int ngrp = half >= nt ? 1 : nt / half;
  const int items = half * ngrp;
  for (int s = 0; s < ns; ++s) {
    const int8_t *sg = a.sig + (size_t)chain * a.T * N2 + (size_t)(nt0 + s * step - 1) * N2;
    kin_tab_build(tab, a.lat, sg, km.nflag, 4 * half, tid, nt);
    for (int cc = 0; cc < ncol; ++cc)
      for (int r = tid; r < N; r += nt) tile[cc * N + r] = kin_multi_pre<R>(tile[cc * N + r], km, c);
    __syncthreads();
    for (int st = 0; st < 8; ++st) {
      const int *t = tab + kin_group_of_stage(st) * half;
      for (int it = tid; it < items; it += nt) {
        const int k = it % half, cg = it / half;
        const int e = t[k];
        const int p1 = e & 2047, p2 = (e >> 11) & 2047;
        const double C = (e >> 22) ? Cp : Cm, S = (e >> 22) ? Sp : Sm;
        for (int cc = cg; cc < ncol; cc += ngrp) {
          double x1 = tile[cc * N + p1], x2 = tile[cc * N + p2];
          double v1 = C * x1 + S * x2;
          double v2 = S * x1 + C * x2;
          tile[cc * N + p1] = v1;
          tile[cc * N + p2] = v2;
        }
      }
      __syncthreads();
    }
    for (int cc = 0; cc < ncol; ++cc)
      for (int r = tid; r < N; r += nt) tile[cc * N + r] = (double)kin_epilogue<R>(tile[cc * N + r], km, c);
  }
  for (int cc = 0; cc < ncol; ++cc)
    for (int r = tid; r < N; r += nt) M[r + (size_t)(col0 + cc) * N] = (R)tile[cc * N + r];
}

template <typename R, int RB, int NC = 0>
__global__ void __launch_bounds__(256) k_kin_cols_multi(KinArgs<R> a, int nt0, int ns, int step) {
  extern __shared__ unsigned char smem_raw[];
  double *tile = reinterpret_cast<double *>(smem_raw);              // [N][RB]
  int *tab = reinterpret_cast<int *>(tile + (size_t)RB * a.lat.N);  // [4][N/2]
  const int N = NC ? NC : a.lat.N, N2 = 2 * N, half = N / 2, nt = blockDim.x, tid = threadIdx.x;
  const int chain = blockIdx.y;
  const KinMat<R> km = a.m[blockIdx.z];
  R *M = km.mat + (size_t)chain * N * N;
  const int row0 = blockIdx.x * RB;
  const int nrow = min(RB, N - row0);
  const ChainConst c = a.cc[chain];
  const int r = tid % RB, cstep = nt / RB;
  if constexpr (NC > 0) {  // the reads of a thread all in flight
    constexpr int CSTEP = KIN_NT / RB, CPT = (NC + CSTEP - 1) / CSTEP;
    R raw[CPT];
    const int rr = row0 + min(r, nrow - 1);
#pragma unroll
    for (int j = 0; j < CPT; ++j) raw[j] = M[rr + (size_t)min(tid / RB + j * CSTEP, NC - 1) * NC];
    if (r < nrow)
#pragma unroll
      for (int j = 0; j < CPT; ++j) {
        const int col = tid / RB + j * CSTEP;
        if (col < NC) tile[col * RB + r] = (double)raw[j];
      }
  } else if (r < nrow) {
    for (int col = tid / RB; col < N; col += cstep) tile[col * RB + r] = (double)M[(row0 + r) + (size_t)col * N];
  }
  const double Cp = c.Cp, Cm = c.Cm, Sp = c.Sp, Sm = c.Sm;
  for (int s = 0; s < ns; ++s) {
    const int8_t *sg = a.sig + (size_t)chain * a.T * N2 + (size_t)(nt0 + s * step - 1) * N2;
    kin_tab_build(tab, a.lat, sg, km.nflag, 4 * half, tid, nt);
    if (r < nrow)
      for (int col = tid / RB; col < N; col += cstep) tile[col * RB + r] = kin_multi_pre<R>(tile[col * RB + r], km, c);
    __syncthreads();
    for (int st = 0; st < 8; ++st) {
      const int *t = tab + kin_group_of_stage(st) * half;
      if (r < nrow)
        for (int k = tid / RB; k < half; k += cstep) {
          const int e = t[k];
          const int p1 = e & 2047, p2 = (e >> 11) & 2047;
          const double C = (e >> 22) ? Cp : Cm, S = (e >> 22) ? Sp : Sm;
          double x1 = tile[p1 * RB + r], x2 = tile[p2 * RB + r];
          double u1 = x1 * C + x2 * S;
          double u2 = x1 * S + x2 * C;
          tile[p1 * RB + r] = u1;
          tile[p2 * RB + r] = u2;
        }
      __syncthreads();
    }
    if (r < nrow)
      for (int col = tid / RB; col < N; col += cstep)
        tile[col * RB + r] = (double)kin_epilogue<R>(tile[col * RB + r], km, c);
  }
  if (r < nrow)
    for (int col = tid / RB; col < N; col += cstep) M[(row0 + r) + (size_t)col * N] = (R)tile[col * RB + r];
}
