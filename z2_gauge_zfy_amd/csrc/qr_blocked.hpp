// Blocked Householder QR and explicit Q for the UDV stabilization, one
// workgroup per chain matrix, for N up to a few hundred.
//
// The reference factors with ZGEQP3 (column pivoting by downdated norms,
// src/stabilization.f90:56).  Here the columns are ordered once by decreasing
// norm ("pre-pivoting") and the ordered matrix gets an unpivoted blocked QR:
// A P = Q R with the same output convention (R rows scaled to unit diagonal
// magnitude, D = |R_ii|, 1-based column permutation, Householder vectors below
// the diagonal, tau).  Pre-pivoting keeps the graded factorization stable for
// the UDV products of DQMC, and lets every trailing update be a block-reflector
// application (compact WY, I - V T V^T) on the FP64 matrix cores instead of one
// pass over the trailing matrix per column.
//
// Panel: NB columns factorized in LDS (one barrier per column, wavefront
// reductions through DPP).  T by the ZLARFT recurrence from the Gram matrix
// V^T V, held in the registers of one wavefront.  Trailing columns: tiles of tc
// columns staged in LDS, W = V^T C, W = T^T W, C -= V W, the two products with
// v_mfma_f64_16x16x4f64.  LDS leading dimensions are odd so that the 16 columns
// an MFMA operand spans fall in different banks.
#pragma once
#include "z2_common.hpp"

struct QrbArgs {
  double *A;     // in: matrix; out: R (normalised rows) + reflectors; chain stride N*N
  double *W;     // workspace, chain stride N*N
  double *tau;   // chain stride N
  int *ipvt;     // chain stride N (1-based)
  double *D;     // chain stride N
  const double *Ain;  // orgqr: reflectors (chain stride N*N)
  const double *tin;  // orgqr: tau
  double *Q;          // orgqr: output (chain stride N*N)
  int N, tc, nopiv;
};

typedef double qrb_d4 __attribute__((ext_vector_type(4)));

// LDS bytes of the blocked kernels for (N, nb, tc): panel and tile (ld N+1),
// T and W (ld nb+1), column norms, tau/beta of the panel
__host__ __device__ inline size_t qrb_lds_bytes(int N, int nb, int tc) {
  const size_t ld = (size_t)N + 1, lw = (size_t)nb + 1;
  size_t d = ld * nb + ld * tc + lw * nb + lw * (tc > nb ? tc : nb) + N + 2 * (size_t)nb;
  return sizeof(double) * d;
}

struct QrbLds {
  int ld, lw;   // leading dimensions: panel/tile (N+1), T/W (nb+1)
  double *Pn;   // panel, column-major, rows relative to the panel top
  double *Ct;   // trailing tile (also the column permutation during the ordering)
  double *Tm;   // nb x nb block-reflector factor T
  double *W1;   // nb x max(tc, nb): V^T C, and the Gram matrix V^T V while T is built
  double *nrm;  // N
  double *s_tau, *s_beta;  // nb
};

__device__ __forceinline__ QrbLds qrb_carve(unsigned char *raw, int N, int nb, int tc) {
  QrbLds L;
  L.ld = N + 1;
  L.lw = nb + 1;
  double *p = reinterpret_cast<double *>(raw);
  L.Pn = p; p += (size_t)L.ld * nb;
  L.Ct = p; p += (size_t)L.ld * tc;
  L.Tm = p; p += (size_t)L.lw * nb;
  L.W1 = p; p += (size_t)L.lw * (tc > nb ? tc : nb);
  L.nrm = p; p += N;
  L.s_tau = p; p += nb;
  L.s_beta = p;
  return L;
}

// element (r, c) of the unit lower-trapezoidal reflector block held in Pn (m rows, w columns)
__device__ __forceinline__ double qrb_v(const double *Pn, int ld, int m, int w, int r, int c) {
  if (r >= m || c >= w || r < c) return 0.0;
  return r == c ? 1.0 : Pn[(size_t)c * ld + r];
}

template <int CTRL>
__device__ __forceinline__ double qrb_dpp(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double qrb_readlane(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}

// sum over the 64 lanes of a wavefront, result in every lane (all lanes active)
__device__ __forceinline__ double qrb_wave_sum(double v) {
  v += qrb_dpp<0xB1>(v);   // quad_perm(1,0,3,2)
  v += qrb_dpp<0x4E>(v);   // quad_perm(2,3,0,1)
  v += qrb_dpp<0x141>(v);  // row_half_mirror
  v += qrb_dpp<0x140>(v);  // row_mirror: every lane holds its 16-lane row sum
  return (qrb_readlane(v, 0) + qrb_readlane(v, 16)) + (qrb_readlane(v, 32) + qrb_readlane(v, 48));
}

// Householder factorization of the panel Pn (m x w) in LDS: reflectors below
// the diagonal, alpha kept on it (beta in s_beta), tau in s_tau.  Column c is
// owned by wave c % nw; a wave updates its columns together.
template <int NB>
__device__ void qrb_panel(QrbLds &L, int m, int w, int last_row) {
  constexpr int NW = 8, CPW = NB / NW;  // 512 threads
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ld = L.ld;
  for (int k = 0; k < w; ++k) {
    if (wave == k % NW) {
      double *ck = L.Pn + (size_t)k * ld;
      double s = 0.0;
      for (int r = k + 1 + lane; r < m; r += 64) s += ck[r] * ck[r];
      s = qrb_wave_sum(s);
      const double xnorm = sqrt(s), alpha = ck[k];
      double tau, beta, scale;
      if (xnorm == 0.0 || k == last_row) {
        tau = 0.0;
        beta = alpha;
        scale = 1.0;
      } else {
        const double xa = fabs(alpha), w_ = fmax(xa, xnorm), z_ = fmin(xa, xnorm);
        const double q = z_ / w_;
        const double nrm = w_ * sqrt(1.0 + q * q);
        beta = -copysign(nrm, alpha);
        tau = (beta - alpha) / beta;
        scale = 1.0 / (alpha - beta);
      }
      for (int r = k + 1 + lane; r < m; r += 64) ck[r] = ck[r] * scale;
      if (lane == 0) {
        L.s_tau[k] = tau;
        L.s_beta[k] = beta;
      }
    }
    __syncthreads();
    const double tk = L.s_tau[k];
    if (tk != 0.0) {
      const double *vk = L.Pn + (size_t)k * ld;
      double d[CPW];
#pragma unroll
      for (int q = 0; q < CPW; ++q) {
        const int c = wave + q * NW;
        d[q] = 0.0;
        if (c > k && c < w) {
          const double *cc = L.Pn + (size_t)c * ld;
          for (int r = k + lane; r < m; r += 64) d[q] += (r == k ? 1.0 : vk[r]) * cc[r];
        }
      }
#pragma unroll
      for (int q = 0; q < CPW; ++q) d[q] = qrb_wave_sum(d[q]);
#pragma unroll
      for (int q = 0; q < CPW; ++q) {
        const int c = wave + q * NW;
        if (c > k && c < w) {
          double *cc = L.Pn + (size_t)c * ld;
          const double t = -tk * d[q];
          for (int r = k + lane; r < m; r += 64) cc[r] = cc[r] + (r == k ? 1.0 : vk[r]) * t;
        }
      }
    }
  }
  __syncthreads();
}

// out (w x ncol, ld L.lw) = V^T B, B = the tile Ct (bv = 0) or V itself (bv = 1).
// One wavefront per 16x16 output block, two accumulators over K.
__device__ void qrb_vt_times(QrbLds &L, int m, int w, int ncol, int bv, double *out) {
  constexpr int NW = 8;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int nbi = (w + 15) >> 4, nbj = (ncol + 15) >> 4;
  const int ld = L.ld;
  for (int blk = wave; blk < nbi * nbj; blk += NW) {
    const int bi = blk % nbi, bj = blk / nbi;
    const int col = 16 * bj + li;
    const int kbeg = 16 * bi;  // V is zero above its diagonal
    qrb_d4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
    for (int kk = kbeg; kk < m; kk += 8) {
      const int r0 = kk + lk, r1 = kk + 4 + lk;
      const double a0 = qrb_v(L.Pn, ld, m, w, r0, 16 * bi + li);
      const double a1 = qrb_v(L.Pn, ld, m, w, r1, 16 * bi + li);
      double b0, b1;
      if (bv) {
        b0 = qrb_v(L.Pn, ld, m, w, r0, col);
        b1 = qrb_v(L.Pn, ld, m, w, r1, col);
      } else {
        b0 = (r0 < m && col < ncol) ? L.Ct[(size_t)col * ld + r0] : 0.0;
        b1 = (r1 < m && col < ncol) ? L.Ct[(size_t)col * ld + r1] : 0.0;
      }
      acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc1, 0, 0, 0);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int ri = 16 * bi + 4 * q + lk, cj = 16 * bj + li;
      out[ri + (size_t)cj * L.lw] = acc0[q] + acc1[q];
    }
  }
}

// T (w x w, upper triangular) of the panel's block reflector by the ZLARFT
// recurrence T[0:i, i] = -tau_i T[0:i, 0:i] (V^T v_i)[0:i], T[i][i] = tau_i.
// Wavefront 0 holds row j of T and of the Gram matrix in lane j's registers.
template <int NB>
__device__ void qrb_build_t(QrbLds &L, int m, int w) {
  qrb_vt_times(L, m, w, w, 1, L.W1);  // Gram
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (wave == 0) {
    const int j = lane;
    double U[NB], T[NB];
#pragma unroll
    for (int l = 0; l < NB; ++l) {
      U[l] = (j < w && l < w && l > j) ? L.W1[j + (size_t)l * L.lw] : 0.0;
      T[l] = 0.0;
    }
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      if (i < w) {
        const double ti = L.s_tau[i];
        double s = 0.0;
#pragma unroll
        for (int l = 0; l < i; ++l) s += T[l] * qrb_readlane(U[i], l);
        T[i] = (j < i) ? -ti * s : (j == i ? ti : 0.0);
      }
    }
    if (j < NB) {
#pragma unroll
      for (int l = 0; l < NB; ++l) L.Tm[j + (size_t)l * L.lw] = (j < w && l < w) ? T[l] : 0.0;
    }
  }
  __syncthreads();
}

// C (m x tcn tile in L.Ct) <- (I - V op(T) V^T) C, op(T) = T^T (trans = 1) or T
template <int NB>
__device__ void qrb_apply_tile(QrbLds &L, int m, int w, int tcn, int trans) {
  constexpr int NW = 8;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int ld = L.ld, lw = L.lw;
  qrb_vt_times(L, m, w, tcn, 0, L.W1);
  __syncthreads();
  // W1 <- op(T) W1: entries e = tid + 512 u (u < 2 covers NB x tc <= 1024),
  // computed into registers, written back after everyone has read W1
  double v[2] = {0.0, 0.0};
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int e = threadIdx.x + 512 * u, i = e % NB, j = e / NB;
    if (j < tcn && i < w) {
      if (trans) {
#pragma unroll
        for (int l = 0; l < NB; ++l)
          if (l <= i) v[u] += L.Tm[l + (size_t)i * lw] * L.W1[l + (size_t)j * lw];
      } else {
#pragma unroll
        for (int l = 0; l < NB; ++l)
          if (l >= i && l < w) v[u] += L.Tm[i + (size_t)l * lw] * L.W1[l + (size_t)j * lw];
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int e = threadIdx.x + 512 * u, i = e % NB, j = e / NB;
    if (j < tcn) L.W1[i + (size_t)j * lw] = v[u];
  }
  __syncthreads();
  // C -= V W1
  const int nbm = (m + 15) >> 4, nbj = (tcn + 15) >> 4;
  for (int blk = wave; blk < nbm * nbj; blk += NW) {
    const int bm = blk % nbm, bj = blk / nbm;
    const int col = 16 * bj + li;
    qrb_d4 acc;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r = 16 * bm + 4 * q + lk;
      acc[q] = (r < m && col < tcn) ? L.Ct[(size_t)col * ld + r] : 0.0;
    }
#pragma unroll
    for (int kk = 0; kk < NB; kk += 4) {
      if (kk < w) {
        const int k = kk + lk;
        const double a = -qrb_v(L.Pn, ld, m, w, 16 * bm + li, k);
        const double b = (k < w && col < tcn) ? L.W1[k + (size_t)col * lw] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r = 16 * bm + 4 * q + lk;
      if (r < m && col < tcn) L.Ct[(size_t)col * ld + r] = acc[q];
    }
  }
  __syncthreads();
}

template <int NB>
__global__ void __launch_bounds__(512) k_qr_blk(QrbArgs a) {
  const int N = a.N, tc = a.tc;
  const int chain = blockIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  double *A = a.A + (size_t)chain * N * N;
  double *W = a.W + (size_t)chain * N * N;
  double *tau = a.tau + (size_t)chain * N;
  int *ipvt = a.ipvt + (size_t)chain * N;
  double *D = a.D + (size_t)chain * N;
  extern __shared__ unsigned char smem_raw[];
  QrbLds L = qrb_carve(smem_raw, N, NB, tc);
  int *perm = reinterpret_cast<int *>(L.Ct);
  // column order: decreasing norm, ties by index
  for (int j = wave; j < N; j += nw) {
    const double *col = A + (size_t)j * N;
    double s = 0.0;
    for (int r = lane; r < N; r += 64) s += col[r] * col[r];
    s = qrb_wave_sum(s);
    if (lane == 0) L.nrm[j] = s;
  }
  __syncthreads();
  for (int j = threadIdx.x; j < N; j += blockDim.x) {
    int rank = j;
    if (!a.nopiv) {
      const double x = L.nrm[j];
      rank = 0;
#pragma unroll 8
      for (int i = 0; i < N; ++i) {
        const double y = L.nrm[i];
        rank += (y > x) || (y == x && i < j);
      }
    }
    perm[rank] = j;
  }
  __syncthreads();
  for (int j = threadIdx.x; j < N; j += blockDim.x) ipvt[j] = perm[j] + 1;
  for (int c = wave; c < N; c += nw) {
    const double *src = A + (size_t)perm[c] * N;
    double *dst = W + (size_t)c * N;
    for (int r = lane; r < N; r += 64) dst[r] = src[r];
  }
  __syncthreads();
  for (int j0 = 0; j0 < N; j0 += NB) {
    const int w = min(NB, N - j0), m = N - j0;
    for (int idx = threadIdx.x; idx < m * w; idx += blockDim.x) {
      const int c = idx / m, r = idx - c * m;
      L.Pn[(size_t)c * L.ld + r] = W[(j0 + r) + (size_t)(j0 + c) * N];
    }
    __syncthreads();
    qrb_panel<NB>(L, m, w, N - 1 - j0);
    if (j0 + w < N) {
      qrb_build_t<NB>(L, m, w);
      for (int c0 = j0 + w; c0 < N; c0 += tc) {
        const int tcn = min(tc, N - c0);
        for (int idx = threadIdx.x; idx < m * tcn; idx += blockDim.x) {
          const int c = idx / m, r = idx - c * m;
          L.Ct[(size_t)c * L.ld + r] = W[(j0 + r) + (size_t)(c0 + c) * N];
        }
        __syncthreads();
        qrb_apply_tile<NB>(L, m, w, tcn, 1);
        for (int idx = threadIdx.x; idx < m * tcn; idx += blockDim.x) {
          const int c = idx / m, r = idx - c * m;
          W[(j0 + r) + (size_t)(c0 + c) * N] = L.Ct[(size_t)c * L.ld + r];
        }
        __syncthreads();
      }
    }
    // panel back: beta on the diagonal, reflectors below
    for (int idx = threadIdx.x; idx < m * w; idx += blockDim.x) {
      const int c = idx / m, r = idx - c * m;
      W[(j0 + r) + (size_t)(j0 + c) * N] = (r == c) ? L.s_beta[c] : L.Pn[(size_t)c * L.ld + r];
    }
    for (int c = threadIdx.x; c < w; c += blockDim.x) tau[j0 + c] = L.s_tau[c];
    __syncthreads();
  }
  // D = |R_ii|; R rows normalised; reflectors copied
  for (int r = threadIdx.x; r < N; r += blockDim.x) {
    const double d = fabs(W[r + (size_t)r * N]);
    D[r] = d;
    L.nrm[r] = 1.0 / d;
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < N * N; idx += blockDim.x) {
    const int j = idx / N, r = idx - j * N;
    const double x = W[idx];
    A[idx] = (r <= j) ? x * L.nrm[r] : x;
  }
}

// Q = H_0 H_1 ... H_{N-1} (ZUNGQR) by block reflectors from the last panel back.
template <int NB>
__global__ void __launch_bounds__(512) k_orgqr_blk(QrbArgs a) {
  const int N = a.N, tc = a.tc;
  const int chain = blockIdx.x;
  const double *A = a.Ain + (size_t)chain * N * N;
  const double *tau = a.tin + (size_t)chain * N;
  double *Q = a.Q + (size_t)chain * N * N;
  extern __shared__ unsigned char smem_raw[];
  QrbLds L = qrb_carve(smem_raw, N, NB, tc);
  const int npan = (N + NB - 1) / NB;
  for (int p = npan - 1; p >= 0; --p) {
    const int j0 = p * NB, w = min(NB, N - j0), m = N - j0;
    for (int idx = threadIdx.x; idx < m * w; idx += blockDim.x) {
      const int c = idx / m, r = idx - c * m;
      L.Pn[(size_t)c * L.ld + r] = A[(j0 + r) + (size_t)(j0 + c) * N];
    }
    for (int c = threadIdx.x; c < w; c += blockDim.x) L.s_tau[c] = tau[j0 + c];
    __syncthreads();
    qrb_build_t<NB>(L, m, w);
    for (int c0 = j0; c0 < N; c0 += tc) {
      const int tcn = min(tc, N - c0);
      // Q[j0:, c0:c0+tcn]: the panels after this one wrote Q[j0+w:, j0+w:];
      // everything else still is the unit matrix
      for (int idx = threadIdx.x; idx < m * tcn; idx += blockDim.x) {
        const int c = idx / m, r = idx - c * m;
        const int gr = j0 + r, gc = c0 + c;
        const bool done = gr >= j0 + w && gc >= j0 + w;
        L.Ct[(size_t)c * L.ld + r] = done ? Q[gr + (size_t)gc * N] : (gr == gc ? 1.0 : 0.0);
      }
      __syncthreads();
      qrb_apply_tile<NB>(L, m, w, tcn, 0);
      for (int idx = threadIdx.x; idx < m * tcn; idx += blockDim.x) {
        const int c = idx / m, r = idx - c * m;
        Q[(j0 + r) + (size_t)(c0 + c) * N] = L.Ct[(size_t)c * L.ld + r];
      }
      __syncthreads();
    }
  }
  // the first panel's update covered all of Q
}
