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
// Panel: NB columns factorized in registers (one barrier per column, wavefront
// reductions through DPP), then stored in LDS.  T by recursive doubling over the
// Gram matrix V^T V (the recursive form of ZLARFT) in LDS.  Trailing columns:
// tiles of tc columns staged in LDS, W = V^T C, W = T^T W, C -= V W, the two
// products with v_mfma_f64_16x16x4f64.  LDS leading dimensions are odd so that the 16 columns
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
  const double *Bq;   // Q^T B: B (chain stride N*N), read as B^T when btrans
  double *Xq;         // Q^T B: output X (chain stride N*N), written as X^T when xtrans
  int btrans, xtrans;
  int N, tc, nopiv;
  int p0w;  // k_qr_trail: panel 0 is in columns 0 .. NB-1 of W (k_qr_panel0), not yet in A
};

typedef double qrb_d4 __attribute__((ext_vector_type(4)));

// LDS bytes of the blocked kernels for (N, nb, tc): panel and tile (ld N+1),
// T and W (ld nb+1), column norms, tau/beta of the panel
__host__ __device__ inline size_t qrb_lds_bytes(int N, int nb, int tc) {
  const size_t ld = (size_t)N + 1, lw = (size_t)nb + 1;
  size_t d = ld * nb + ld * tc + lw * nb + lw * (tc > nb ? tc : nb) + N + 2 * (size_t)nb + 2 * ld;
  return sizeof(double) * d + sizeof(int) * (size_t)N;
}

struct QrbLds {
  int ld, lw;   // leading dimensions: panel/tile (N+1), T/W (nb+1)
  double *Pn;   // panel, column-major, rows relative to the panel top
  double *Ct;   // trailing tile (also the column permutation during the ordering)
  double *Tm;   // nb x nb block-reflector factor T
  double *W1;   // nb x max(tc, nb): V^T C, and the Gram matrix V^T V while T is built
  double *nrm;  // N
  double *s_tau, *s_beta;  // nb
  double *vb;   // 2 x ld: reflector of the current panel column (double-buffered)
  int *perm;    // N: column order
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
  L.s_beta = p; p += nb;
  L.vb = p; p += 2 * (size_t)L.ld;
  L.perm = reinterpret_cast<int *>(p);
  return L;
}

// element (r, c) of the unit lower-trapezoidal reflector block held in Pn (m rows,
// w columns); the LDS load is unconditional (clamped address) so that no branch
// separates consecutive loads
__device__ __forceinline__ double qrb_v(const double *Pn, int ld, int m, int w, int r, int c) {
  const double x = Pn[(size_t)min(c, w - 1) * ld + min(r, m - 1)];
  return (r >= m || c >= w || r < c) ? 0.0 : (r == c ? 1.0 : x);
}

// p[i] if ok else 0, with an unconditional load of p[ok ? i : 0]
__device__ __forceinline__ double qrb_ld(const double *p, size_t i, bool ok) {
  const double x = p[ok ? i : 0];
  return ok ? x : 0.0;
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

// DPP move into the rows selected by ROWMASK, 0 elsewhere
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double qrb_dpp_rows(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWMASK, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWMASK, 0xF, false);
  return __hiloint2double(hi, lo);
}

// sum over the 64 lanes of a wavefront, result in every lane (all lanes active):
// row sums by quad permutes and mirrors, then row_bcast:15 / row_bcast:31 carry the
// row sums up to lane 63, which is read once
__device__ __forceinline__ double qrb_wave_sum(double v) {
  v += qrb_dpp<0xB1>(v);   // quad_perm(1,0,3,2)
  v += qrb_dpp<0x4E>(v);   // quad_perm(2,3,0,1)
  v += qrb_dpp<0x141>(v);  // row_half_mirror
  v += qrb_dpp<0x140>(v);  // row_mirror: every lane holds its 16-lane row sum
  v += qrb_dpp_rows<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
  v += qrb_dpp_rows<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3: lane 63 has the total
  return qrb_readlane(v, 63);
}

// C independent wavefront sums by the tree of qrb_wave_sum, stage by stage, so
// that their latencies overlap (each sum is bit-identical to qrb_wave_sum's)
template <int C>
__device__ __forceinline__ void qrb_wave_sum_n(double (&v)[C]) {
#pragma unroll
  for (int q = 0; q < C; ++q) v[q] += qrb_dpp<0xB1>(v[q]);
#pragma unroll
  for (int q = 0; q < C; ++q) v[q] += qrb_dpp<0x4E>(v[q]);
#pragma unroll
  for (int q = 0; q < C; ++q) v[q] += qrb_dpp<0x141>(v[q]);
#pragma unroll
  for (int q = 0; q < C; ++q) v[q] += qrb_dpp<0x140>(v[q]);
#pragma unroll
  for (int q = 0; q < C; ++q) v[q] += qrb_dpp_rows<0x142, 0xA>(v[q]);
#pragma unroll
  for (int q = 0; q < C; ++q) v[q] += qrb_dpp_rows<0x143, 0xC>(v[q]);
#pragma unroll
  for (int q = 0; q < C; ++q) v[q] = qrb_readlane(v[q], 63);
}

// The sums of v[Q..C-1] only (the same trees as qrb_wave_sum_n), for Q = qa, the
// first entry still needed; qa >= C: none
template <int C, int Q>
__device__ __forceinline__ void qrb_wave_sum_tail(double (&v)[C]) {
#pragma unroll
  for (int q = Q; q < C; ++q) v[q] += qrb_dpp<0xB1>(v[q]);
#pragma unroll
  for (int q = Q; q < C; ++q) v[q] += qrb_dpp<0x4E>(v[q]);
#pragma unroll
  for (int q = Q; q < C; ++q) v[q] += qrb_dpp<0x141>(v[q]);
#pragma unroll
  for (int q = Q; q < C; ++q) v[q] += qrb_dpp<0x140>(v[q]);
#pragma unroll
  for (int q = Q; q < C; ++q) v[q] += qrb_dpp_rows<0x142, 0xA>(v[q]);
#pragma unroll
  for (int q = Q; q < C; ++q) v[q] += qrb_dpp_rows<0x143, 0xC>(v[q]);
#pragma unroll
  for (int q = Q; q < C; ++q) v[q] = qrb_readlane(v[q], 63);
}

template <int C, int Q = 0>
__device__ __forceinline__ void qrb_wave_sum_from(double (&v)[C], int qa) {
  if constexpr (Q < C) {
    if (qa == Q) {
      qrb_wave_sum_tail<C, Q>(v);
      return;
    }
    qrb_wave_sum_from<C, Q + 1>(v, qa);
  }
}

// Householder factorization of a panel (m x w, w <= NB) in registers: reflectors
// below the diagonal, alpha kept on it (beta in s_beta), tau in s_tau.  Wavefront w owns columns
// c = w + 8 q, lane l rows r = l + 64 t (RPL = ceil(N / 64)).  One barrier per
// column; the reflector reaches the other wavefronts through LDS (double-buffered).
// Each wavefront applies reflector k to all its trailing columns with one
// interleaved set of wavefront sums, and the owner of column k+1 then computes
// reflector k+1 before the barrier.
// Reads the panel (m x w) from global memory: column c at Wp + col(c) * ldW with
// col(c) = cperm[c] (or c); leaves the factored panel in L.Pn.
template <int NB, int RPL>
__device__ __forceinline__ void qrb_reflector(QrbLds &L, double (&x)[NB / 8][RPL], int k, int m, int last_row,
                                              int lane) {
#pragma clang fp contract(fast)
  constexpr int NW = 8, CPW = NB / NW;
  const int qk = k / NW;
  double col[RPL];
#pragma unroll
  for (int t = 0; t < RPL; ++t) {
    col[t] = x[0][t];
#pragma unroll
    for (int q = 1; q < CPW; ++q)
      if (q == qk) col[t] = x[q][t];
  }
  double s = 0.0, ak = 0.0;
#pragma unroll
  for (int t = 0; t < RPL; ++t) {
    const int r = lane + 64 * t;
    if (r > k && r < m) s += col[t] * col[t];
    if (t == (k >> 6)) ak = col[t];
  }
  s = qrb_wave_sum(s);
  const double alpha = qrb_readlane(ak, k & 63);
  double tau, beta, scale;
  if (s == 0.0 || k == last_row) {
    tau = 0.0;
    beta = alpha;
    scale = 1.0;
  } else {
    // |(alpha, x)|: one square root while alpha^2 + |x|^2 is far from overflow and
    // underflow (the column step's latency is this scalar chain), else the scaled
    // form w sqrt(1 + (z/w)^2)
    const double a2 = alpha * alpha + s;
    double nrm;
    if (a2 > 1e-290 && a2 < 1e290) {
      nrm = sqrt(a2);
    } else {
      const double xnorm = sqrt(s), xa = fabs(alpha), w_ = fmax(xa, xnorm), z_ = fmin(xa, xnorm);
      const double q_ = z_ / w_;
      nrm = w_ * sqrt(1.0 + q_ * q_);
    }
    beta = -copysign(nrm, alpha);
    tau = (beta - alpha) / beta;
    scale = 1.0 / (alpha - beta);
  }
  double *vb = L.vb + (size_t)(k & 1) * L.ld;
#pragma unroll
  for (int t = 0; t < RPL; ++t) {
    const int r = lane + 64 * t;
    if (r > k && r < m) col[t] = col[t] * scale;
    if (r < m) vb[r] = (r < k) ? 0.0 : (r == k ? 1.0 : col[t]);
  }
#pragma unroll
  for (int q = 0; q < CPW; ++q)
    if (q == qk)
#pragma unroll
      for (int t = 0; t < RPL; ++t) x[q][t] = col[t];
  if (lane == 0) {
    L.s_tau[k] = tau;
    L.s_beta[k] = beta;
  }
}

template <int NB, int RPL>
__device__ void qrb_panel_reg(QrbLds &L, const double *Wp, int ldW, const int *cperm, int m, int w,
                              int last_row) {
#pragma clang fp contract(fast)
  constexpr int NW = 8, CPW = NB / NW;  // 512 threads
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double x[CPW][RPL];
#pragma unroll
  for (int q = 0; q < CPW; ++q) {
    const int c = wave + NW * q;
    const double *src = Wp + (size_t)(c < w ? (cperm ? cperm[c] : c) : 0) * ldW;
#pragma unroll
    for (int t = 0; t < RPL; ++t) {
      const int r = lane + 64 * t;
      x[q][t] = qrb_ld(src, (size_t)r, c < w && r < m);
    }
  }
  if (wave == 0) qrb_reflector<NB, RPL>(L, x, 0, m, last_row, lane);
  __syncthreads();
  for (int k = 0; k < w; ++k) {
    const double tk = L.s_tau[k];
    const double *vb = L.vb + (size_t)(k & 1) * L.ld;
    double vr[RPL];
#pragma unroll
    for (int t = 0; t < RPL; ++t) {
      const int r = lane + 64 * t;
      vr[t] = qrb_ld(vb, (size_t)r, r < m);
    }
    // this wavefront's columns c = wave + NW q still to be updated: q >= qa (c > k);
    // the sums of the others are skipped (uniform per wavefront)
    const int qa = k < wave ? 0 : (k - wave) / NW + 1;
    if (tk != 0.0 && qa < CPW) {
      double d[CPW];
#pragma unroll
      for (int q = 0; q < CPW; ++q) {
        d[q] = 0.0;
        const int c = wave + NW * q;
        if (c > k && c < w)
#pragma unroll
          for (int t = 0; t < RPL; ++t) d[q] += vr[t] * x[q][t];
      }
      qrb_wave_sum_from<CPW>(d, qa);
#pragma unroll
      for (int q = 0; q < CPW; ++q) {
        const int c = wave + NW * q;
        if (c > k && c < w) {
          const double tt = -tk * d[q];
#pragma unroll
          for (int t = 0; t < RPL; ++t) x[q][t] = x[q][t] + vr[t] * tt;
        }
      }
    }
    const int k1 = k + 1;
    if (k1 < w && wave == k1 % NW) qrb_reflector<NB, RPL>(L, x, k1, m, last_row, lane);
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < CPW; ++q)
#pragma unroll
    for (int t = 0; t < RPL; ++t) {
      const int c = wave + NW * q, r = lane + 64 * t;
      if (c < w && r < m) L.Pn[(size_t)c * L.ld + r] = x[q][t];
    }
  __syncthreads();
}

// global <-> LDS tile transfers: wavefront w moves columns w, w + 8, ..., lane l rows
// l + 64 t; all loads of a thread are issued before its LDS stores
template <int RPL>
__device__ __forceinline__ void qrb_tile_load(double *dst, int ldd, const double *src, size_t lds, const int *cperm,
                                              int m, int ncol) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int c0 = wave; c0 < ncol; c0 += 32) {
    double v[4][RPL];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int c = c0 + 8 * u;
      const double *s = src + (size_t)(c < ncol ? (cperm ? cperm[c] : c) : 0) * lds;
#pragma unroll
      for (int t = 0; t < RPL; ++t) {
        const int r = lane + 64 * t;
        v[u][t] = qrb_ld(s, (size_t)r, c < ncol && r < m);
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int c = c0 + 8 * u;
#pragma unroll
      for (int t = 0; t < RPL; ++t) {
        const int r = lane + 64 * t;
        if (c < ncol && r < m) dst[(size_t)c * ldd + r] = v[u][t];
      }
    }
  }
}

// out (w x ncol, ld L.lw) = V^T B, B = the tile Ct (bv = 0) or V itself (bv = 1).
// One wavefront per 16x16 output block, two accumulators over K; with at most
// half as many blocks as wavefronts, two wavefronts share a block, each over half
// of K, and the second half's sum is added through out (one more barrier).
__device__ void qrb_vt_times(QrbLds &L, int m, int w, int ncol, int bv, double *out) {
  constexpr int NW = 8;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int nbi = (w + 15) >> 4, nbj = (ncol + 15) >> 4, nb = nbi * nbj;
  const int ld = L.ld;
  const bool split = 2 * nb <= NW;
  for (int blk = split ? wave % nb : wave; blk < nb && (!split || wave < 2 * nb); blk += NW) {
    const int bi = blk % nbi, bj = blk / nbi;
    const int col = 16 * bj + li;
    const int kbeg = 16 * bi;  // V is zero above its diagonal
    int klo = kbeg, khi = m;
    if (split) {
      const int steps = (m - kbeg + 15) >> 4, hs = (steps + 1) >> 1;
      if (wave < nb) khi = min(m, kbeg + 16 * hs);
      else klo = kbeg + 16 * hs;
    }
    qrb_d4 acc[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) acc[u] = (qrb_d4){0.0, 0.0, 0.0, 0.0};
    for (int kk = klo; kk < khi; kk += 16) {
      double av[4], bw[4];  // four K steps of operands loaded together, four chains
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int r = kk + 4 * u + lk;
        av[u] = qrb_v(L.Pn, ld, m, w, r, 16 * bi + li);
        bw[u] = bv ? qrb_v(L.Pn, ld, m, w, r, col) : qrb_ld(L.Ct, (size_t)col * ld + r, r < m && col < ncol);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u], bw[u], acc[u], 0, 0, 0);
    }
    if (split) {
      // (rows beyond m in the second half read as zero through qrb_v / qrb_ld)
      if (wave >= nb)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int ri = 16 * bi + 4 * q + lk, cj = 16 * bj + li;
          out[ri + (size_t)cj * L.lw] = (acc[0][q] + acc[1][q]) + (acc[2][q] + acc[3][q]);
        }
      __syncthreads();
      if (wave < nb)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int ri = 16 * bi + 4 * q + lk, cj = 16 * bj + li;
          const double lo = (acc[0][q] + acc[1][q]) + (acc[2][q] + acc[3][q]);
          out[ri + (size_t)cj * L.lw] = lo + out[ri + (size_t)cj * L.lw];
        }
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int ri = 16 * bi + 4 * q + lk, cj = 16 * bj + li;
        out[ri + (size_t)cj * L.lw] = (acc[0][q] + acc[1][q]) + (acc[2][q] + acc[3][q]);
      }
    }
  }
  if (split && wave >= 2 * nb) __syncthreads();  // the idle wavefronts meet the barrier above
}

// One doubling level of the block reflector factor: every pair of H x H diagonal
// blocks (T1 at rows/columns o..o+H-1, T2 at o+H..o+2H-1) is merged into
// [T1, -T1 (V1^T V2) T2; 0, T2].  V1^T V2 is the Gram block above the diagonal
// of W1 (entries of reflectors >= w read as zero); Y = (V1^T V2) T2 goes to the
// transposed positions below the diagonal of W1, which no later level reads.
template <int NB, int H>
__device__ __forceinline__ void qrb_t_level(QrbLds &L, int w) {
  constexpr int B = 2 * H, CNT = (NB / B) * H * H;
  const int lw = L.lw;
  for (int e = threadIdx.x; e < CNT; e += blockDim.x) {
    const int blk = e / (H * H), rr = (e / H) % H, cc = e % H;  // constant divisors
    const int o = blk * B, r = o + rr, c = o + H + cc;
    double y = 0.0;
#pragma unroll
    for (int l = 0; l < H; ++l) {  // T2 is upper triangular: zero for l > cc
      const int g = o + H + l;
      const double G = (r < w && g < w) ? L.W1[r + (size_t)g * lw] : 0.0;
      y += G * L.Tm[g + (size_t)c * lw];
    }
    L.W1[c + (size_t)r * lw] = y;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < CNT; e += blockDim.x) {
    const int blk = e / (H * H), rr = (e / H) % H, cc = e % H;
    const int o = blk * B, r = o + rr, c = o + H + cc;
    double x = 0.0;
#pragma unroll
    for (int l = 0; l < H; ++l)  // T1 is upper triangular: zero for l < rr
      x += L.Tm[r + (size_t)(o + l) * lw] * L.W1[c + (size_t)(o + l) * lw];
    L.Tm[r + (size_t)c * lw] = -x;
  }
  __syncthreads();
}

// T (NB x NB, upper triangular, zero at indices >= w) of the panel's block
// reflector H_0 ... H_{w-1} = I - V T V^T: diagonal tau_i, then log2(NB)
// doubling levels over the Gram matrix V^T V (the recursive form of ZLARFT).
template <int NB>
__device__ void qrb_build_t(QrbLds &L, int m, int w) {
  qrb_vt_times(L, m, w, w, 1, L.W1);  // Gram
  for (int e = threadIdx.x; e < NB * NB; e += blockDim.x) {
    const int i = e % NB, l = e / NB;
    L.Tm[i + (size_t)l * L.lw] = (i == l && i < w) ? L.s_tau[i] : 0.0;
  }
  __syncthreads();
  qrb_t_level<NB, 1>(L, w);
  qrb_t_level<NB, 2>(L, w);
  qrb_t_level<NB, 4>(L, w);
  qrb_t_level<NB, 8>(L, w);
  if constexpr (NB >= 32) qrb_t_level<NB, 16>(L, w);
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
  // W1 <- op(T) W1 on the matrix cores: wavefront b < (NB/16) x ceil(tcn/16)
  // computes output block b over K = lmax (T is stored with its zeros: triangle,
  // rows/columns >= w; W1 has rows up to lmax), written back after everyone has
  // read W1.  Result item q of lane l is block element (4q + l/16, l%16).
  const int lmax = 16 * ((w + 15) >> 4);
  constexpr int TBI = NB / 16;
  const int tbj = (tcn + 15) >> 4;
  const bool tw = wave < TBI * tbj;
  qrb_d4 tacc = (qrb_d4){0.0, 0.0, 0.0, 0.0};
  if (tw) {
    const int bi = wave % TBI, bj = wave / TBI;
    const int i = 16 * bi + li, j = min(16 * bj + li, tcn - 1);
    for (int l0 = 0; l0 < lmax; l0 += 4) {
      const int l = l0 + lk;
      const double av = trans ? L.Tm[l + (size_t)i * lw] : L.Tm[i + (size_t)l * lw];
      const double bv = L.W1[l + (size_t)j * lw];
      tacc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, tacc, 0, 0, 0);
    }
  }
  __syncthreads();
  if (tw) {
    const int bi = wave % TBI, bj = wave / TBI, col = 16 * bj + li;
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (col < tcn) L.W1[(16 * bi + 4 * q + lk) + (size_t)col * lw] = tacc[q];
  }
  __syncthreads();
  // C -= V W1: a wavefront takes row blocks wave and wave + 8 (then + 16, ...) and
  // both column blocks, so its accumulators are independent chains sharing operands
  const int nbm = (m + 15) >> 4, nbj = (tcn + 15) >> 4;
  for (int bm0 = wave; bm0 < nbm; bm0 += NW * 2) {
    qrb_d4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int bm = bm0 + NW * i, col = 16 * j + li;
        const bool ok = bm < nbm && j < nbj && col < tcn;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = 16 * bm + 4 * q + lk;
          acc[i][j][q] = qrb_ld(L.Ct, (size_t)col * ld + r, ok && r < m);
        }
      }
#pragma unroll
    for (int u = 0; u < NB / 4; ++u) {
      if (4 * u < w) {
        const int k = 4 * u + lk;
        double av[2], bw[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) av[i] = -qrb_v(L.Pn, ld, m, w, 16 * (bm0 + NW * i) + li, k);
#pragma unroll
        for (int j = 0; j < 2; ++j) bw[j] = qrb_ld(L.W1, k + (size_t)(16 * j + li) * lw, k < w && 16 * j + li < tcn);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
            if (bm0 + NW * i < nbm && j < nbj)
              acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i], bw[j], acc[i][j], 0, 0, 0);
      }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int bm = bm0 + NW * i, col = 16 * j + li;
        if (bm < nbm && j < nbj && col < tcn)
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int r = 16 * bm + 4 * q + lk;
            if (r < m) L.Ct[(size_t)col * ld + r] = acc[i][j][q];
          }
      }
  }
  __syncthreads();
}

template <int NB, int RPL>
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
  // column order: decreasing norm, ties by index (norms of 4 columns per wavefront at once)
  for (int j0 = wave; j0 < N; j0 += 4 * nw) {
    double sq[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int j = j0 + nw * u;
      const double *col = A + (size_t)(j < N ? j : 0) * N;
      double v[RPL];
#pragma unroll
      for (int t = 0; t < RPL; ++t) {
        const int r = lane + 64 * t;
        v[t] = (j < N && r < N) ? col[r] : 0.0;
      }
      sq[u] = 0.0;
#pragma unroll
      for (int t = 0; t < RPL; ++t) sq[u] += v[t] * v[t];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) sq[u] = qrb_wave_sum(sq[u]);
    if (lane == 0)
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (j0 + nw * u < N) L.nrm[j0 + nw * u] = sq[u];
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
    L.perm[rank] = j;
  }
  __syncthreads();
  for (int j = threadIdx.x; j < N; j += blockDim.x) ipvt[j] = L.perm[j] + 1;
  // panels; the first one reads the columns of A through the order, later ones W
  for (int j0 = 0; j0 < N; j0 += NB) {
    const int w = min(NB, N - j0), m = N - j0;
    const bool first = j0 == 0;
    qrb_panel_reg<NB, RPL>(L, first ? A : W + j0 + (size_t)j0 * N, N, first ? L.perm : nullptr, m, w,
                           N - 1 - j0);
    if (j0 + w < N) {
      qrb_build_t<NB>(L, m, w);
      for (int c0 = j0 + w; c0 < N; c0 += tc) {
        const int tcn = min(tc, N - c0);
        if (first) qrb_tile_load<RPL>(L.Ct, L.ld, A, N, L.perm + c0, m, tcn);
        else qrb_tile_load<RPL>(L.Ct, L.ld, W + j0 + (size_t)c0 * N, N, nullptr, m, tcn);
        __syncthreads();
        qrb_apply_tile<NB>(L, m, w, tcn, 1);
        // rows of this panel are final R rows: normalised into A; the rest back to W.
        // (The first panel still reads A's columns through the order: all to W.)
        for (int c = wave; c < tcn; c += nw)
#pragma unroll
          for (int t = 0; t < RPL; ++t) {
            const int r = lane + 64 * t;
            if (r < m) {
              const double x = L.Ct[(size_t)c * L.ld + r];
              if (r < w && !first) A[(j0 + r) + (size_t)(c0 + c) * N] = x / fabs(L.s_beta[r]);
              else W[(j0 + r) + (size_t)(c0 + c) * N] = x;
            }
          }
        __syncthreads();
      }
      if (first)  // A is consumed: the first panel's R rows of the trailing columns
        for (int c = w + wave; c < N; c += nw)
          for (int r = lane; r < w; r += 64) A[r + (size_t)c * N] = W[r + (size_t)c * N] / fabs(L.s_beta[r]);
    }
    // the panel's columns of A: R rows above (already in A), diagonal block and reflectors
    for (int c = wave; c < w; c += nw)
#pragma unroll
      for (int t = 0; t < RPL; ++t) {
        const int r = lane + 64 * t;
        if (r < m) {
          double x;
          if (r < c) x = L.Pn[(size_t)c * L.ld + r] / fabs(L.s_beta[r]);
          else if (r == c) x = L.s_beta[c] / fabs(L.s_beta[c]);
          else x = L.Pn[(size_t)c * L.ld + r];
          A[(j0 + r) + (size_t)(j0 + c) * N] = x;
        }
      }
    for (int c = threadIdx.x; c < w; c += blockDim.x) {
      tau[j0 + c] = L.s_tau[c];
      D[j0 + c] = fabs(L.s_beta[c]);
    }
    __syncthreads();
  }
}

// Q = H_0 H_1 ... H_{N-1} (ZUNGQR) by block reflectors from the last panel back.
template <int NB, int RPL>
__global__ void __launch_bounds__(512) k_orgqr_blk(QrbArgs a) {
  const int N = a.N, tc = a.tc;
  const int chain = blockIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const double *A = a.Ain + (size_t)chain * N * N;
  const double *tau = a.tin + (size_t)chain * N;
  double *Q = a.Q + (size_t)chain * N * N;
  extern __shared__ unsigned char smem_raw[];
  QrbLds L = qrb_carve(smem_raw, N, NB, tc);
  const int npan = (N + NB - 1) / NB;
  for (int p = npan - 1; p >= 0; --p) {
    const int j0 = p * NB, w = min(NB, N - j0), m = N - j0;
    qrb_tile_load<RPL>(L.Pn, L.ld, A + j0 + (size_t)j0 * N, N, nullptr, m, w);
    for (int c = threadIdx.x; c < w; c += blockDim.x) L.s_tau[c] = tau[j0 + c];
    __syncthreads();
    qrb_build_t<NB>(L, m, w);
    for (int c0 = j0; c0 < N; c0 += tc) {
      const int tcn = min(tc, N - c0);
      // Q[j0:, c0:c0+tcn]: the panels after this one wrote Q[j0+w:, j0+w:];
      // everything else still is the unit matrix
      if (c0 >= j0 + w) {
        qrb_tile_load<RPL>(L.Ct, L.ld, Q + j0 + (size_t)c0 * N, N, nullptr, m, tcn);
        for (int c = wave; c < tcn; c += nw)
          for (int r = lane; r < w && r < m; r += 64) L.Ct[(size_t)c * L.ld + r] = 0.0;
      } else {
        for (int c = wave; c < tcn; c += nw)
#pragma unroll
          for (int t = 0; t < RPL; ++t) {
            const int r = lane + 64 * t;
            if (r < m) L.Ct[(size_t)c * L.ld + r] = (j0 + r == c0 + c) ? 1.0 : 0.0;
          }
      }
      __syncthreads();
      qrb_apply_tile<NB>(L, m, w, tcn, 0);
      for (int c = wave; c < tcn; c += nw)
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
          const int r = lane + 64 * t;
          if (r < m) Q[(j0 + r) + (size_t)(c0 + c) * N] = L.Ct[(size_t)c * L.ld + r];
        }
      __syncthreads();
    }
  }
  // the first panel's update covered all of Q
}

// Explicit Q spread over many workgroups per matrix: k_orgqr_t builds the T factor
// of every panel (one workgroup per panel and matrix) into Tws, then k_orgqr_tiles
// applies the block reflectors to each column tile.  Tws: per matrix npan x NB x NB.
template <int NB, int RPL>
__global__ void __launch_bounds__(512) k_orgqr_t(QrbArgs a, double *Tws) {
  const int N = a.N, tc = a.tc, p = blockIdx.x, chain = blockIdx.y;
  const int npan = (N + NB - 1) / NB;
  const double *A = a.Ain + (size_t)chain * N * N;
  const double *tau = a.tin + (size_t)chain * N;
  extern __shared__ unsigned char smem_raw[];
  QrbLds L = qrb_carve(smem_raw, N, NB, tc);
  const int j0 = p * NB, w = min(NB, N - j0), m = N - j0;
  qrb_tile_load<RPL>(L.Pn, L.ld, A + j0 + (size_t)j0 * N, N, nullptr, m, w);
  for (int c = threadIdx.x; c < w; c += blockDim.x) L.s_tau[c] = tau[j0 + c];
  __syncthreads();
  qrb_build_t<NB>(L, m, w);
  double *T = Tws + ((size_t)chain * npan + p) * NB * NB;
  for (int e = threadIdx.x; e < NB * NB; e += blockDim.x) T[e] = L.Tm[(e % NB) + (size_t)(e / NB) * L.lw];
}

// One column tile of Q per workgroup, all panels applied to it in one launch:
// Q[:, tile] = H_0 (H_1 (... (H_last e_tile))); the tile (all N rows) stays in LDS,
// panel p's block reflector touches rows j0.. of it, and panels that start after
// the tile's last column leave it unchanged.  The tiles with the most panels are
// dispatched first.  Each panel application is the same as in k_orgqr_blk.
// X = Q^T op(B) (or its transpose) with Q held as the block reflectors of the QR
// (T factors of every panel in Tws, from k_orgqr_t): one column tile of op(B) per
// workgroup, kept in LDS while panel 0, 1, ... apply (I - V T^T V^T) to its rows
// j0.. (Q^T = H_{n-1} ... H_0) -- the trailing update of the factorization,
// without forming Q.  op(B) = B^T is read along B's rows, X^T written along X's rows.
template <int NB, int RPL>
__global__ void __launch_bounds__(512) k_qt_apply(QrbArgs a, const double *Tws) {
  const int N = a.N, tc = a.tc, chain = blockIdx.y;
  const int npan = (N + NB - 1) / NB;
  const int c0 = blockIdx.x * tc, tcn = min(tc, N - c0);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const double *A = a.Ain + (size_t)chain * N * N;
  const double *B = a.Bq + (size_t)chain * N * N;
  double *X = a.Xq + (size_t)chain * N * N;
  extern __shared__ unsigned char smem_raw[];
  QrbLds L = qrb_carve(smem_raw, N, NB, tc);
  if (a.btrans) {
    for (int e = threadIdx.x; e < 32 * N; e += blockDim.x) {  // tc <= 32
      const int c = e & 31, r = e >> 5;
      if (c < tcn) L.Ct[(size_t)c * L.ld + r] = B[(c0 + c) + (size_t)r * N];
    }
  } else {
    qrb_tile_load<RPL>(L.Ct, L.ld, B + (size_t)c0 * N, N, nullptr, N, tcn);
  }
  for (int p = 0; p < npan; ++p) {
    const int j0 = p * NB, w = min(NB, N - j0), m = N - j0;
    __syncthreads();  // the tile is staged / the previous panel is done with Pn and Tm
    qrb_tile_load<RPL>(L.Pn, L.ld, A + j0 + (size_t)j0 * N, N, nullptr, m, w);
    const double *T = Tws + ((size_t)chain * npan + p) * NB * NB;
    for (int e = threadIdx.x; e < NB * NB; e += blockDim.x) L.Tm[(e % NB) + (size_t)(e / NB) * L.lw] = T[e];
    __syncthreads();
    QrbLds Lp = L;
    Lp.Ct = L.Ct + j0;  // rows j0.. of the tile
    qrb_apply_tile<NB>(Lp, m, w, tcn, 1);
  }
  __syncthreads();
  if (a.xtrans) {
    for (int e = threadIdx.x; e < 32 * N; e += blockDim.x) {
      const int c = e & 31, r = e >> 5;
      if (c < tcn) X[(c0 + c) + (size_t)r * N] = L.Ct[(size_t)c * L.ld + r];
    }
  } else {
    for (int c = wave; c < tcn; c += nw)
#pragma unroll
      for (int t = 0; t < RPL; ++t) {
        const int r = lane + 64 * t;
        if (r < N) X[r + (size_t)(c0 + c) * N] = L.Ct[(size_t)c * L.ld + r];
      }
  }
}

template <int NB, int RPL>
__global__ void __launch_bounds__(512) k_orgqr_tiles(QrbArgs a, const double *Tws) {
  const int N = a.N, tc = a.tc, chain = blockIdx.y;
  const int npan = (N + NB - 1) / NB, ntile = (N + tc - 1) / tc;
  const int c0 = (ntile - 1 - (int)blockIdx.x) * tc, tcn = min(tc, N - c0);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const double *A = a.Ain + (size_t)chain * N * N;
  double *Q = a.Q + (size_t)chain * N * N;
  extern __shared__ unsigned char smem_raw[];
  QrbLds L = qrb_carve(smem_raw, N, NB, tc);
  for (int c = wave; c < tcn; c += nw)
#pragma unroll
    for (int t = 0; t < RPL; ++t) {
      const int r = lane + 64 * t;
      if (r < N) L.Ct[(size_t)c * L.ld + r] = (r == c0 + c) ? 1.0 : 0.0;
    }
  const int plast = min(npan - 1, (c0 + tcn - 1) / NB);
  for (int p = plast; p >= 0; --p) {
    const int j0 = p * NB, w = min(NB, N - j0), m = N - j0;
    __syncthreads();  // the previous panel's application is done with Pn and Tm
    qrb_tile_load<RPL>(L.Pn, L.ld, A + j0 + (size_t)j0 * N, N, nullptr, m, w);
    const double *T = Tws + ((size_t)chain * npan + p) * NB * NB;
    for (int e = threadIdx.x; e < NB * NB; e += blockDim.x) L.Tm[(e % NB) + (size_t)(e / NB) * L.lw] = T[e];
    __syncthreads();
    QrbLds Lp = L;
    Lp.Ct = L.Ct + j0;  // rows j0.. of the tile
    qrb_apply_tile<NB>(Lp, m, w, tcn, 0);
  }
  __syncthreads();
  for (int c = wave; c < tcn; c += nw)
#pragma unroll
    for (int t = 0; t < RPL; ++t) {
      const int r = lane + 64 * t;
      if (r < N) Q[r + (size_t)(c0 + c) * N] = L.Ct[(size_t)c * L.ld + r];
    }
}

// The QR spread over many workgroups per matrix.  k_qr_order: column norms, the
// order (decreasing norm, ties by index) into ipvt.  k_qr_panel0 factors panel 0
// through that order while its other workgroups copy the remaining ordered
// columns into W (for N <= NB: the copy, then k_qr_panel).  Per later panel,
// k_qr_panel (one workgroup per matrix) factors the panel from W (or the
// look-ahead below does), writes its columns of A (normalised R rows of the
// diagonal block, signs, reflectors), tau, D and the block reflector's T into
// Tws; k_qr_trail applies it to the trailing columns with one workgroup per
// column tile: the panel's rows become final normalised R rows in A, the rest go
// back to W; with tiles as wide as panels, the workgroup of the first tile then
// factors the next panel (look-ahead), so a panel step is one launch.  Each tile
// sees the same operations as in k_qr_blk.
template <int RPL>
__global__ void __launch_bounds__(512) k_qr_order(QrbArgs a) {
  const int N = a.N, chain = blockIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const double *A = a.A + (size_t)chain * N * N;
  int *ipvt = a.ipvt + (size_t)chain * N;
  extern __shared__ unsigned char smem_raw[];
  double *nrm = reinterpret_cast<double *>(smem_raw);
  for (int j0 = wave; j0 < N; j0 += 4 * nw) {
    double sq[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int j = j0 + nw * u;
      const double *col = A + (size_t)(j < N ? j : 0) * N;
      double v[RPL];
#pragma unroll
      for (int t = 0; t < RPL; ++t) {
        const int r = lane + 64 * t;
        v[t] = (j < N && r < N) ? col[r] : 0.0;
      }
      sq[u] = 0.0;
#pragma unroll
      for (int t = 0; t < RPL; ++t) sq[u] += v[t] * v[t];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) sq[u] = qrb_wave_sum(sq[u]);
    if (lane == 0)
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (j0 + nw * u < N) nrm[j0 + nw * u] = sq[u];
  }
  __syncthreads();
  for (int j = threadIdx.x; j < N; j += blockDim.x) {
    int rank = j;
    if (!a.nopiv) {
      const double x = nrm[j];
      rank = 0;
#pragma unroll 8
      for (int i = 0; i < N; ++i) {
        const double y = nrm[i];
        rank += (y > x) || (y == x && i < j);
      }
    }
    ipvt[rank] = j + 1;
  }
}

// Panel p of one matrix: factored from the m x w block at Wp (column stride ldW,
// global or LDS; column c at Wp + cperm[c] ldW if cperm) into L.Pn; writes its
// columns of A (normalised R rows of the diagonal block, signs, reflectors; of
// Out instead when given), tau, D and its T into Tws.
template <int NB, int RPL>
__device__ void qrb_panel_step(const QrbArgs &a, QrbLds &L, double *Tws, int chain, int p, const double *Wp,
                               int ldW, const int *cperm = nullptr, double *Out = nullptr) {
  const int N = a.N;
  const int npan = (N + NB - 1) / NB;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  double *A = Out ? Out : a.A + (size_t)chain * N * N;
  double *tau = a.tau + (size_t)chain * N;
  double *D = a.D + (size_t)chain * N;
  const int j0 = p * NB, w = min(NB, N - j0), m = N - j0;
  qrb_panel_reg<NB, RPL>(L, Wp, ldW, cperm, m, w, N - 1 - j0);
  // T of every panel, the last one too: the explicit Q and the Q^T application
  // that follow the factorization use them as they are
  qrb_build_t<NB>(L, m, w);
  double *T = Tws + ((size_t)chain * npan + p) * NB * NB;
  for (int e = threadIdx.x; e < NB * NB; e += blockDim.x) T[e] = L.Tm[(e % NB) + (size_t)(e / NB) * L.lw];
  for (int c = wave; c < w; c += nw)
#pragma unroll
    for (int t = 0; t < RPL; ++t) {
      const int r = lane + 64 * t;
      if (r < m) {
        double x;
        if (r < c) x = L.Pn[(size_t)c * L.ld + r] / fabs(L.s_beta[r]);
        else if (r == c) x = L.s_beta[c] / fabs(L.s_beta[c]);
        else x = L.Pn[(size_t)c * L.ld + r];
        A[(j0 + r) + (size_t)(j0 + c) * N] = x;
      }
    }
  for (int c = threadIdx.x; c < w; c += blockDim.x) {
    tau[j0 + c] = L.s_tau[c];
    D[j0 + c] = fabs(L.s_beta[c]);
  }
}

template <int NB, int RPL>
__global__ void __launch_bounds__(512) k_qr_panel(QrbArgs a, double *Tws, int p) {
  const int N = a.N, chain = blockIdx.x;
  extern __shared__ unsigned char smem_raw[];
  QrbLds L = qrb_carve(smem_raw, N, NB, a.tc);
  const int j0 = p * NB;
  qrb_panel_step<NB, RPL>(a, L, Tws, chain, p, a.W + (size_t)chain * N * N + j0 + (size_t)j0 * N, N);
}

// Panel 0 and the column order in one launch (N > NB): workgroup (0, chain)
// factors panel 0 from A through the order of k_qr_order (ipvt) into columns
// 0 .. NB-1 of W, the other workgroups of the chain copy the ordered columns NB ..
// N-1 of A into W (W = A P beyond the panel).  A is only read here, so it may be
// written from the next launch on, which takes panel 0 from W and copies it to A.
template <int NB, int RPL>
__global__ void __launch_bounds__(512) k_qr_panel0(QrbArgs a, double *Tws) {
  const int N = a.N, chain = blockIdx.y;
  const int *ipvt = a.ipvt + (size_t)chain * N;
  const double *A = a.A + (size_t)chain * N * N;
  double *W = a.W + (size_t)chain * N * N;
  if (blockIdx.x == 0) {
    extern __shared__ unsigned char smem_raw[];
    QrbLds L = qrb_carve(smem_raw, N, NB, a.tc);
    for (int c = threadIdx.x; c < NB; c += blockDim.x) L.perm[c] = ipvt[c] - 1;
    __syncthreads();
    qrb_panel_step<NB, RPL>(a, L, Tws, chain, 0, A, N, L.perm, W);
    return;
  }
  for (int c = NB + (int)blockIdx.x - 1; c < N; c += (int)gridDim.x - 1) {
    const double *src = A + (size_t)(ipvt[c] - 1) * N;
    for (int r = threadIdx.x; r < N; r += blockDim.x) W[r + (size_t)c * N] = src[r];
  }
}

// lookahead (tile width == NB): the workgroup of tile 0, whose columns are the
// next panel's, goes on to factor panel p + 1 from its updated tile in LDS while
// the other tiles are updated (the next launch is then panel p + 1's trail).
template <int NB, int RPL>
__global__ void __launch_bounds__(512) k_qr_trail(QrbArgs a, double *Tws, int p, int lookahead) {
  const int N = a.N, tc = a.tc, chain = blockIdx.y;
  const int npan = (N + NB - 1) / NB;
  const int j0 = p * NB, w = min(NB, N - j0), m = N - j0;
  const int c0 = j0 + w + blockIdx.x * tc;
  if (c0 >= N) return;  // whole workgroup
  const int tcn = min(tc, N - c0);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  double *A = a.A + (size_t)chain * N * N;
  double *W = a.W + (size_t)chain * N * N;
  const double *D = a.D + (size_t)chain * N;
  extern __shared__ unsigned char smem_raw[];
  QrbLds L = qrb_carve(smem_raw, N, NB, tc);
  // reflectors below the diagonal (panel 0: from W, where k_qr_panel0 left it)
  const double *Vp = p == 0 && a.p0w ? W : A + j0 + (size_t)j0 * N;
  qrb_tile_load<RPL>(L.Pn, L.ld, Vp, N, nullptr, m, w);
  const double *T = Tws + ((size_t)chain * npan + p) * NB * NB;
  for (int e = threadIdx.x; e < NB * NB; e += blockDim.x) L.Tm[(e % NB) + (size_t)(e / NB) * L.lw] = T[e];
  qrb_tile_load<RPL>(L.Ct, L.ld, W + j0 + (size_t)c0 * N, N, nullptr, m, tcn);
  __syncthreads();
  if (p == 0 && a.p0w && blockIdx.x == 0)  // panel 0's columns of A (m = N rows)
    for (int c = wave; c < w; c += nw)
#pragma unroll
      for (int t = 0; t < RPL; ++t) {
        const int r = lane + 64 * t;
        if (r < m) A[r + (size_t)c * N] = L.Pn[(size_t)c * L.ld + r];
      }
  qrb_apply_tile<NB>(L, m, w, tcn, 1);
  for (int c = wave; c < tcn; c += nw)
#pragma unroll
    for (int t = 0; t < RPL; ++t) {
      const int r = lane + 64 * t;
      if (r < m) {
        const double x = L.Ct[(size_t)c * L.ld + r];
        if (r < w) A[(j0 + r) + (size_t)(c0 + c) * N] = x / D[j0 + r];
        else W[(j0 + r) + (size_t)(c0 + c) * N] = x;
      }
    }
  if (lookahead && blockIdx.x == 0) {
    __syncthreads();  // V of panel p (Pn), T and W1 are free; the tile stays in Ct
    qrb_panel_step<NB, RPL>(a, L, Tws, chain, p + 1, L.Ct + w, L.ld);
  }
}

