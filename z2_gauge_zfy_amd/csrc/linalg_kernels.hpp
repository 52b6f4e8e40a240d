// FP64 dense kernels for the UDV stabilization (reference src/stabilization.f90:
// 42-200): column-pivoted Householder QR (ZGEQP3 semantics, ZLAQP2 column
// norm downdating), explicit Q (ZUNGQR), batched GEMM on the FP64 matrix cores,
// triangular solves (ZTRSM) and the permutations (ZLAPMR/ZLAPMT).
// All matrices N x N column-major, one per chain (stride N*N).
#pragma once
#include "z2_common.hpp"

// ------------------------------------------------------------------ QRP
// One workgroup per matrix, matrix in global memory (L2 resident).
// out: A holds R (rows normalised by |R_ii|, src/stabilization.f90:61-66) in the
// upper triangle and the Householder vectors below; tau; ipvt (1-based); D = |R_ii|.
struct QrpArgs {
  double *A;
  double *tau;     // chain stride N
  int *ipvt;       // chain stride N
  double *D;       // chain stride N (may alias a propagator D vector)
  int N;
  int nopiv;       // 1: plain Householder QR (no column exchanges)
};

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__global__ void __launch_bounds__(1024) k_qrp(QrpArgs a) {
  const int N = a.N;
  const int chain = blockIdx.x;
  double *A = a.A + (size_t)chain * N * N;
  double *tau = a.tau + (size_t)chain * N;
  int *ipvt = a.ipvt + (size_t)chain * N;
  extern __shared__ unsigned char smem_raw[];
  double *vn1 = reinterpret_cast<double *>(smem_raw);
  double *vn2 = vn1 + N;
  double *redv = vn2 + N;                          // 32
  int *redi = reinterpret_cast<int *>(redv + 32);  // 32
  __shared__ double s_tau, s_beta, s_scale;
  __shared__ int s_pvt;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const double tol3z = sqrt(1.1102230246251565e-16);

  // initial column norms
  for (int j = wid; j < N; j += nw) {
    const double *col = A + (size_t)j * N;
    double s = 0.0;
    for (int r = lane; r < N; r += 64) s += col[r] * col[r];
    s = wave_sum(s);
    if (lane == 0) {
      vn1[j] = sqrt(s);
      vn2[j] = vn1[j];
      ipvt[j] = j + 1;
    }
  }
  __syncthreads();
  for (int i = 0; i < N; ++i) {
    // pivot: first index of the largest vn1[i..N)
    double best = -1.0;
    int bi = N;
    for (int j = i + threadIdx.x; j < N; j += blockDim.x) {
      double v = vn1[j];
      if (v > best) { best = v; bi = j; }
    }
    for (int off = 32; off > 0; off >>= 1) {
      double ov = __shfl_xor(best, off, 64);
      int oi = __shfl_xor(bi, off, 64);
      if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
    }
    if (lane == 0) { redv[wid] = best; redi[wid] = bi; }
    __syncthreads();
    if (threadIdx.x == 0) {
      double bv = redv[0];
      int bj = redi[0];
      for (int w = 1; w < nw; ++w)
        if (redv[w] > bv || (redv[w] == bv && redi[w] < bj)) { bv = redv[w]; bj = redi[w]; }
      s_pvt = bj;
    }
    __syncthreads();
    const int pvt = a.nopiv ? i : s_pvt;
    if (pvt != i) {
      double *cp = A + (size_t)pvt * N, *ci = A + (size_t)i * N;
      for (int r = threadIdx.x; r < N; r += blockDim.x) {
        double t = cp[r];
        cp[r] = ci[r];
        ci[r] = t;
      }
      if (threadIdx.x == 0) {
        int it = ipvt[pvt];
        ipvt[pvt] = ipvt[i];
        ipvt[i] = it;
        vn1[pvt] = vn1[i];
        vn2[pvt] = vn2[i];
      }
    }
    __syncthreads();
    // Householder reflector for A[i:, i] (dlarfg)
    double *ci = A + (size_t)i * N;
    double s = 0.0;
    for (int r = i + 1 + threadIdx.x; r < N; r += blockDim.x) s += ci[r] * ci[r];
    s = wave_sum(s);
    if (lane == 0) redv[wid] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double ss = 0.0;
      for (int w = 0; w < nw; ++w) ss += redv[w];
      double xnorm = sqrt(ss);
      double alpha = ci[i];
      if (xnorm == 0.0 || i == N - 1) {
        s_tau = 0.0;
        s_beta = alpha;
        s_scale = 1.0;
      } else {
        double xa = fabs(alpha), w_ = fmax(xa, xnorm), z_ = fmin(xa, xnorm);
        double q = z_ / w_;
        double nrm = w_ * sqrt(1.0 + q * q);
        double beta = -copysign(nrm, alpha);
        s_tau = (beta - alpha) / beta;
        s_scale = 1.0 / (alpha - beta);
        s_beta = beta;
      }
    }
    __syncthreads();
    const double tau_i = s_tau, sc = s_scale;
    for (int r = i + 1 + threadIdx.x; r < N; r += blockDim.x) ci[r] = ci[r] * sc;
    __syncthreads();
    if (threadIdx.x == 0) {
      ci[i] = s_beta;
      tau[i] = tau_i;
    }
    // apply H(i) to A[i:, i+1:], one wavefront per column; downdate norms
    for (int j = i + 1 + wid; j < N; j += nw) {
      double *cj = A + (size_t)j * N;
      double w = 0.0;
      if (tau_i != 0.0) {
        for (int r = i + lane; r < N; r += 64) {
          double v = (r == i) ? 1.0 : ci[r];
          w += cj[r] * v;
        }
        w = wave_sum(w);
        double t = -tau_i * w;
        for (int r = i + lane; r < N; r += 64) {
          double v = (r == i) ? 1.0 : ci[r];
          cj[r] = cj[r] + v * t;
        }
      }
      // norm downdate (ZLAQP2)
      double aij = __shfl(cj[i], 0, 64);  // all lanes read after own writes (lane owning row i wrote it)
      if (vn1[j] != 0.0) {
        double rr = fabs(aij) / vn1[j];
        double temp = 1.0 - rr * rr;
        temp = temp < 0.0 ? 0.0 : temp;
        double qq = vn1[j] / vn2[j];
        double temp2 = temp * qq * qq;
        if (temp2 <= tol3z) {
          double s2 = 0.0;
          for (int r = i + 1 + lane; r < N; r += 64) s2 += cj[r] * cj[r];
          s2 = wave_sum(s2);
          if (lane == 0) {
            vn1[j] = (i < N - 1) ? sqrt(s2) : 0.0;
            vn2[j] = vn1[j];
          }
        } else if (lane == 0) {
          vn1[j] = vn1[j] * sqrt(temp);
        }
      }
    }
    __syncthreads();
  }
  // D = |R_ii|, rows of R normalised
  for (int r = threadIdx.x; r < N; r += blockDim.x) a.D[(size_t)chain * N + r] = fabs(A[r + (size_t)r * N]);
  __syncthreads();
  for (int idx = threadIdx.x; idx < N * N; idx += blockDim.x) {
    int j = idx / N, r = idx - j * N;
    if (r <= j) A[idx] = A[idx] / a.D[(size_t)chain * N + r];
  }
}

// Register-blocked variant of k_qrp / k_orgqr for N <= 64*RPL: every lane
// keeps its RPL rows of the Householder vector in registers and a wavefront
// updates CPW columns at once (one load/store round trip per column group).
#define QR_REG_THREADS 512  // 2 waves per SIMD: room for CPW*RPL doubles per lane

// apply H = I - tau v v^T (v in registers vr[t] for rows r = i + lane + 64 t)
// to columns j of M[i:, j]; optional ZLAQP2 norm downdating.
template <int RPL, int CPW>
__device__ __forceinline__ void apply_h_cols(double *M, int N, int i, const double (&vr)[RPL], double tau,
                                             int jbeg, int jstep, double *vn1, double *vn2, int lane) {
  const double tol3z = 1.4901161193847656e-08;  // sqrt(2^-53)
  for (int j0 = jbeg; j0 < N; j0 += CPW * jstep) {
    double x[CPW][RPL];
    double d[CPW];
#pragma unroll
    for (int cc = 0; cc < CPW; ++cc) {
      const int j = j0 + cc * jstep;
      d[cc] = 0.0;
#pragma unroll
      for (int t = 0; t < RPL; ++t) {
        const int r = i + lane + 64 * t;
        x[cc][t] = (j < N && r < N) ? M[r + (size_t)j * N] : 0.0;
      }
    }
    if (tau != 0.0) {
#pragma unroll
      for (int cc = 0; cc < CPW; ++cc)
#pragma unroll
        for (int t = 0; t < RPL; ++t) d[cc] += x[cc][t] * vr[t];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int cc = 0; cc < CPW; ++cc) d[cc] += __shfl_xor(d[cc], off, 64);
#pragma unroll
      for (int cc = 0; cc < CPW; ++cc) {
        const int j = j0 + cc * jstep;
        const double tt = -tau * d[cc];
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
          const int r = i + lane + 64 * t;
          x[cc][t] = x[cc][t] + vr[t] * tt;
          if (j < N && r < N) M[r + (size_t)j * N] = x[cc][t];
        }
      }
    }
    if (vn1) {
#pragma unroll
      for (int cc = 0; cc < CPW; ++cc) {
        const int j = j0 + cc * jstep;
        if (j >= N) continue;
        const double aij = __shfl(x[cc][0], 0, 64);  // row i sits in lane 0, t = 0
        const double v1 = vn1[j];
        if (v1 != 0.0) {
          double rr = fabs(aij) / v1;
          double temp = 1.0 - rr * rr;
          temp = temp < 0.0 ? 0.0 : temp;
          double qq = v1 / vn2[j];
          double temp2 = temp * qq * qq;
          if (temp2 <= tol3z) {
            double s2 = 0.0;
#pragma unroll
            for (int t = 0; t < RPL; ++t) {
              const int r = i + lane + 64 * t;
              if (r > i && r < N) s2 += x[cc][t] * x[cc][t];
            }
            s2 = wave_sum(s2);
            if (lane == 0) {
              vn1[j] = (i < N - 1) ? sqrt(s2) : 0.0;
              vn2[j] = vn1[j];
            }
          } else if (lane == 0) {
            vn1[j] = v1 * sqrt(temp);
          }
        }
      }
    }
  }
}

template <int RPL, int CPW>
__global__ void __launch_bounds__(QR_REG_THREADS) k_qrp_reg(QrpArgs a) {
  const int N = a.N;
  const int chain = blockIdx.x;
  double *A = a.A + (size_t)chain * N * N;
  double *tau = a.tau + (size_t)chain * N;
  int *ipvt = a.ipvt + (size_t)chain * N;
  extern __shared__ unsigned char smem_raw[];
  double *vn1 = reinterpret_cast<double *>(smem_raw);
  double *vn2 = vn1 + N;
  double *redv = vn2 + N;
  int *redi = reinterpret_cast<int *>(redv + 32);
  __shared__ double s_tau, s_beta, s_scale;
  __shared__ int s_pvt;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int j = wid; j < N; j += nw) {
    const double *col = A + (size_t)j * N;
    double s = 0.0;
    for (int r = lane; r < N; r += 64) s += col[r] * col[r];
    s = wave_sum(s);
    if (lane == 0) {
      vn1[j] = sqrt(s);
      vn2[j] = vn1[j];
      ipvt[j] = j + 1;
    }
  }
  __syncthreads();
  for (int i = 0; i < N; ++i) {
    double best = -1.0;
    int bi = N;
    for (int j = i + threadIdx.x; j < N; j += blockDim.x) {
      double v = vn1[j];
      if (v > best) { best = v; bi = j; }
    }
    for (int off = 32; off > 0; off >>= 1) {
      double ov = __shfl_xor(best, off, 64);
      int oi = __shfl_xor(bi, off, 64);
      if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
    }
    if (lane == 0) { redv[wid] = best; redi[wid] = bi; }
    __syncthreads();
    if (threadIdx.x == 0) {
      double bv = redv[0];
      int bj = redi[0];
      for (int w = 1; w < nw; ++w)
        if (redv[w] > bv || (redv[w] == bv && redi[w] < bj)) { bv = redv[w]; bj = redi[w]; }
      s_pvt = bj;
    }
    __syncthreads();
    const int pvt = a.nopiv ? i : s_pvt;
    double *ci = A + (size_t)i * N;
    if (pvt != i) {
      double *cp = A + (size_t)pvt * N;
      for (int r = threadIdx.x; r < N; r += blockDim.x) {
        double t = cp[r];
        cp[r] = ci[r];
        ci[r] = t;
      }
      if (threadIdx.x == 0) {
        int it = ipvt[pvt];
        ipvt[pvt] = ipvt[i];
        ipvt[i] = it;
        vn1[pvt] = vn1[i];
        vn2[pvt] = vn2[i];
      }
      __syncthreads();
    }
    // reflector: wave 0 computes the norm of A[i+1:, i] (dlarfg)
    if (wid == 0) {
      double s = 0.0;
      for (int r = i + 1 + lane; r < N; r += 64) s += ci[r] * ci[r];
      s = wave_sum(s);
      if (lane == 0) {
        double xnorm = sqrt(s);
        double alpha = ci[i];
        if (xnorm == 0.0 || i == N - 1) {
          s_tau = 0.0;
          s_beta = alpha;
          s_scale = 1.0;
        } else {
          double xa = fabs(alpha), w_ = fmax(xa, xnorm), z_ = fmin(xa, xnorm);
          double q = z_ / w_;
          double nrm = w_ * sqrt(1.0 + q * q);
          double beta = -copysign(nrm, alpha);
          s_tau = (beta - alpha) / beta;
          s_scale = 1.0 / (alpha - beta);
          s_beta = beta;
        }
      }
    }
    __syncthreads();
    const double tau_i = s_tau, sc = s_scale;
    double vr[RPL];
#pragma unroll
    for (int t = 0; t < RPL; ++t) {
      const int r = i + lane + 64 * t;
      vr[t] = (r == i) ? 1.0 : (r < N ? ci[r] * sc : 0.0);
    }
    __syncthreads();  // every wave holds v before wave 0 overwrites column i below
    apply_h_cols<RPL, CPW>(A, N, i, vr, tau_i, i + 1 + wid, nw, vn1, vn2, lane);
    // the reflector itself (wave 0 owns the write-back of column i)
    if (wid == 0) {
#pragma unroll
      for (int t = 0; t < RPL; ++t) {
        const int r = i + lane + 64 * t;
        if (r > i && r < N) ci[r] = vr[t];
      }
      if (lane == 0) {
        ci[i] = s_beta;
        tau[i] = tau_i;
      }
    }
    __syncthreads();
  }
  for (int r = threadIdx.x; r < N; r += blockDim.x) a.D[(size_t)chain * N + r] = fabs(A[r + (size_t)r * N]);
  __syncthreads();
  for (int idx = threadIdx.x; idx < N * N; idx += blockDim.x) {
    int j = idx / N, r = idx - j * N;
    if (r <= j) A[idx] = A[idx] / a.D[(size_t)chain * N + r];
  }
}

template <int RPL, int CPW>
__global__ void __launch_bounds__(QR_REG_THREADS) k_orgqr_reg(const double *Ain, const double *tauin, double *Qout,
                                                    int N) {
  const int chain = blockIdx.x;
  const double *A = Ain + (size_t)chain * N * N;
  const double *tau = tauin + (size_t)chain * N;
  double *Q = Qout + (size_t)chain * N * N;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int idx = threadIdx.x; idx < N * N; idx += blockDim.x) {
    int j = idx / N, r = idx - j * N;
    Q[idx] = (r == j) ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int i = N - 1; i >= 0; --i) {
    const double ti = tau[i];
    const double *v = A + (size_t)i * N;
    double vr[RPL];
#pragma unroll
    for (int t = 0; t < RPL; ++t) {
      const int r = i + lane + 64 * t;
      vr[t] = (r == i) ? 1.0 : (r < N ? v[r] : 0.0);
    }
    apply_h_cols<RPL, CPW>(Q, N, i, vr, ti, i + 1 + wid, nw, nullptr, nullptr, lane);
    double *qi = Q + (size_t)i * N;
    for (int r = threadIdx.x; r < N; r += blockDim.x) {
      if (r < i) qi[r] = 0.0;
      else if (r == i) qi[r] = 1.0 - ti;
      else qi[r] = -ti * v[r];
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ ORGQR
// Q (N x N) from the reflectors (below the diagonal of A) and tau; writes Q to
// Q (may not alias A).  ZUNG2R order: H(N-1) applied first to the unit matrix.
__global__ void __launch_bounds__(1024) k_orgqr(const double *Ain, const double *tauin, double *Qout,
                                                int N) {
  const int chain = blockIdx.x;
  const double *A = Ain + (size_t)chain * N * N;
  const double *tau = tauin + (size_t)chain * N;
  double *Q = Qout + (size_t)chain * N * N;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int idx = threadIdx.x; idx < N * N; idx += blockDim.x) {
    int j = idx / N, r = idx - j * N;
    Q[idx] = (r == j) ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int i = N - 1; i >= 0; --i) {
    const double ti = tau[i];
    const double *v = A + (size_t)i * N;  // v[i] = 1 implicit, v[r>i] = A[r,i]
    // columns j > i: Q[i:, j] -= ti * v (v^T Q[i:, j])
    for (int j = i + 1 + wid; j < N; j += nw) {
      double *qj = Q + (size_t)j * N;
      double w = 0.0;
      for (int r = i + lane; r < N; r += 64) w += qj[r] * ((r == i) ? 1.0 : v[r]);
      w = wave_sum(w);
      double t = -ti * w;
      for (int r = i + lane; r < N; r += 64) qj[r] = qj[r] + ((r == i) ? 1.0 : v[r]) * t;
    }
    // column i: (1 - ti) at i, -ti*v below, 0 above
    double *qi = Q + (size_t)i * N;
    for (int r = threadIdx.x; r < N; r += blockDim.x) {
      if (r < i) qi[r] = 0.0;
      else if (r == i) qi[r] = 1.0 - ti;
      else qi[r] = -ti * v[r];
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ GEMM
// C = op(A) op(B) (+ beta C), batched over chains, FP64 MFMA 16x16x4.
// trans flags: 0 = N, 1 = T.  triA / triB: use only the upper triangle
// (row <= col, in stored coordinates) of A / B (ZTRMM as a masked GEMM).
struct GemmArgs {
  const double *A, *B;
  double *C;
  long sA, sB, sC;      // chain strides
  int M, Nc, K;         // C is M x Nc, op(A) M x K, op(B) K x Nc
  int lda, ldb, ldc;
  int ta, tb, triA, triB;
  double alpha, beta;
};

#define GT 64   // tile M, N
#define GK 16   // tile K

// C = alpha op(A) op(B) + beta C, batched over chains (grid.z), FP64 MFMA 16x16x4.
// trans flags: 0 = N, 1 = T.  triA / triB: use only the upper triangle (row <= col,
// in stored coordinates) of A / B (ZTRMM as a masked GEMM).  A 64 x 64 tile of C
// per workgroup (4 wavefronts of 2 x 2 MFMA blocks); each K step of 16 stages
// op(A) and op(B) in LDS with lanes running along each operand's contiguous
// dimension (coalesced for every transpose case), and the next step's values are
// read into registers while the current one is multiplied.
struct GemmTileLd {
  // 4 values per thread per operand and K step: element (i, k) of a GT x GK panel
  int i[4], k[4];
};

// staging coordinates of thread t for a panel whose storage is contiguous along
// the tile dimension (i) or along K
__device__ __forceinline__ GemmTileLd gemm_coords(int t, bool contig_i) {
  GemmTileLd c;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int e = t + 256 * q;  // 0 .. GT*GK-1
    if (contig_i) {
      c.i[q] = e % GT;
      c.k[q] = e / GT;
    } else {
      c.k[q] = e % GK;
      c.i[q] = e / GK;
    }
  }
  return c;
}

// element (r, c) of op(X) (R x Cc), 0 outside or below the diagonal when tri;
// the address is clamped in range and the value selected (no branch around the load)
__device__ __forceinline__ double gemm_ld(const double *X, int ld, int trans, int tri, int r, int c, int R,
                                          int Cc) {
  const bool in = r < R && c < Cc;
  const int sr = trans ? c : r, sc = trans ? r : c;
  const bool keep = in && !(tri && sr > sc);
  const double v = X[in ? sr + (size_t)sc * ld : 0];
  return keep ? v : 0.0;
}

// Two independent products of the same shape in one launch: grid.z = nfirst chains
// of g0, then the chains of g1 (g1 = g0 and grid.z = nfirst for one product).
__global__ void __launch_bounds__(256) k_gemm(GemmArgs g0, GemmArgs g1, int nfirst) {
  __shared__ double As[GK][GT + 1];  // As[k][m]
  __shared__ double Bs[GK][GT + 1];  // Bs[k][n]
  const bool second = (int)blockIdx.z >= nfirst;
  const GemmArgs g = second ? g1 : g0;
  const int chain = (int)blockIdx.z - (second ? nfirst : 0);
  const double *A = g.A + chain * g.sA;
  const double *B = g.B + chain * g.sB;
  double *C = g.C + chain * g.sC;
  const int m0 = blockIdx.x * GT, n0 = blockIdx.y * GT;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int wm = (wid & 1) * 32, wn = (wid >> 1) * 32;  // each wave: 32x32 = 2x2 MFMA tiles
  // op(A)(m, k) is contiguous along m when A is not transposed; op(B)(k, n) is
  // contiguous along n when B is transposed
  const GemmTileLd ca = gemm_coords(threadIdx.x, g.ta == 0);
  const GemmTileLd cb = gemm_coords(threadIdx.x, g.tb != 0);
  typedef double d4 __attribute__((ext_vector_type(4)));
  d4 acc[2][2];
  for (int x = 0; x < 2; ++x)
    for (int y = 0; y < 2; ++y) acc[x][y] = (d4){0.0, 0.0, 0.0, 0.0};
  double ra[4], rb[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    ra[q] = gemm_ld(A, g.lda, g.ta, g.triA, m0 + ca.i[q], ca.k[q], g.M, g.K);
    rb[q] = gemm_ld(B, g.ldb, g.tb, g.triB, cb.k[q], n0 + cb.i[q], g.K, g.Nc);
  }
  for (int k0 = 0; k0 < g.K; k0 += GK) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      As[ca.k[q]][ca.i[q]] = ra[q];
      Bs[cb.k[q]][cb.i[q]] = rb[q];
    }
    __syncthreads();
    if (k0 + GK < g.K) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        ra[q] = gemm_ld(A, g.lda, g.ta, g.triA, m0 + ca.i[q], k0 + GK + ca.k[q], g.M, g.K);
        rb[q] = gemm_ld(B, g.ldb, g.tb, g.triB, k0 + GK + cb.k[q], n0 + cb.i[q], g.K, g.Nc);
      }
    }
#pragma unroll
    for (int kk = 0; kk < GK; kk += 4) {
      const int kl = kk + (lane >> 4), il = lane & 15;
      double av[2], bv[2];
      for (int x = 0; x < 2; ++x) av[x] = As[kl][wm + 16 * x + il];
      for (int y = 0; y < 2; ++y) bv[y] = Bs[kl][wn + 16 * y + il];
      for (int x = 0; x < 2; ++x)
        for (int y = 0; y < 2; ++y)
          acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[x], bv[y], acc[x][y], 0, 0, 0);
    }
    __syncthreads();
  }
  for (int x = 0; x < 2; ++x)
    for (int y = 0; y < 2; ++y)
      for (int q = 0; q < 4; ++q) {
        // v_mfma_f64_16x16x4f64 result: lane l, item q holds D[4q + l/16][l%16]
        int row = m0 + wm + 16 * x + 4 * q + (lane >> 4);
        int col = n0 + wn + 16 * y + (lane & 15);
        if (row < g.M && col < g.Nc) {
          double v = g.alpha * acc[x][y][q];
          size_t o = row + (size_t)col * g.ldc;
          C[o] = (g.beta != 0.0) ? v + g.beta * C[o] : v;
        }
      }
}

// ------------------------------------------------------------------ TRSM
// Solve R x = b for each right-hand side, R upper triangular (upper part of
// Rm incl. diagonal).  Right-hand side c, row i lives at X[i*rs + c*cs].
// mode 0 (ZTRSM 'L','U','N','N'): x_k = b_k / R_kk, then b_i -= x_k R_ik.
// mode 1 (ZTRSM 'R','U','C','N' applied to each row): x_k = (1/R_kk) * b_k, then
//        b_i -= R_ik x_k.
// A workgroup owns cbt (<= CBT) right-hand sides, staged in LDS.
#define CBT 16
__global__ void __launch_bounds__(256) k_trsm(const double *Rm, double *X, int N, int rs, int cs,
                                              int mode, int cbt) {
  extern __shared__ unsigned char smem_raw[];
  double *xb = reinterpret_cast<double *>(smem_raw);  // [cbt][N]
  double *rk = xb + cbt * N;                          // column k of R
  const int chain = blockIdx.y;
  const double *R = Rm + (size_t)chain * N * N;
  double *Xc = X + (size_t)chain * N * N;
  const int c0 = blockIdx.x * cbt;
  const int nc = min(cbt, N - c0);
  for (int idx = threadIdx.x; idx < nc * N; idx += blockDim.x) {
    int c = idx / N, i = idx - c * N;
    xb[c * N + i] = Xc[(size_t)i * rs + (size_t)(c0 + c) * cs];
  }
  __syncthreads();
  for (int k = N - 1; k >= 0; --k) {
    for (int i = threadIdx.x; i <= k; i += blockDim.x) rk[i] = R[i + (size_t)k * N];
    __syncthreads();
    if (threadIdx.x < nc) {
      double *x = xb + threadIdx.x * N;
      if (mode == 0) x[k] = x[k] / rk[k];
      else x[k] = (1.0 / rk[k]) * x[k];
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < nc * k; idx += blockDim.x) {
      int c = idx / k, i = idx - c * k;
      double *x = xb + c * N;
      if (mode == 0) x[i] = x[i] - x[k] * rk[i];
      else x[i] = x[i] - rk[i] * x[k];
    }
    __syncthreads();
  }
  for (int idx = threadIdx.x; idx < nc * N; idx += blockDim.x) {
    int c = idx / N, i = idx - c * N;
    Xc[(size_t)i * rs + (size_t)(c0 + c) * cs] = xb[c * N + i];
  }
}

// Same solve with X in registers: a group of 16 lanes owns one column of X
// (lane l holds rows l + 16 j, j < MJ).  Back substitution by 16-row blocks:
// the block's rows are solved with the pivot divided by its owner lane and
// broadcast by a lane shuffle (no barrier), then the rows above take the
// block's 16 updates.  Every element still receives its updates one by one in
// descending k with one multiply and one subtract each: the arithmetic and its
// order are those of k_trsm (and ZTRSM).  Columns of R are staged in LDS 16 at
// a time.  grid: (ceil(N/16), chains), 256 threads = 16 columns.
// lane kk of every 16-lane row of the wavefront, in all lanes of that row (DPP
// row_newbcast; kk a constant once the caller's loop is unrolled)
template <int KK>
__device__ __forceinline__ double trsm_row_bcast_t(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x150 + KK, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x150 + KK, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double trsm_row_bcast(double v, int kk) {
  switch (kk) {
    case 0: return trsm_row_bcast_t<0>(v);
    case 1: return trsm_row_bcast_t<1>(v);
    case 2: return trsm_row_bcast_t<2>(v);
    case 3: return trsm_row_bcast_t<3>(v);
    case 4: return trsm_row_bcast_t<4>(v);
    case 5: return trsm_row_bcast_t<5>(v);
    case 6: return trsm_row_bcast_t<6>(v);
    case 7: return trsm_row_bcast_t<7>(v);
    case 8: return trsm_row_bcast_t<8>(v);
    case 9: return trsm_row_bcast_t<9>(v);
    case 10: return trsm_row_bcast_t<10>(v);
    case 11: return trsm_row_bcast_t<11>(v);
    case 12: return trsm_row_bcast_t<12>(v);
    case 13: return trsm_row_bcast_t<13>(v);
    case 14: return trsm_row_bcast_t<14>(v);
    default: return trsm_row_bcast_t<15>(v);
  }
}

// dsc (chain stride N), if given: the right-hand side's solve index i divided by
// dsc[i] as it is read (the D^{-1} of the G recompute); perm (chain stride N,
// 1-based), if given: the solution written to Y at solve index perm[i] - 1 instead
// of in place (the P of the G recompute) -- the arithmetic of a division kernel
// before and a permutation kernel after the solve
template <int MJ>
__global__ void __launch_bounds__(256) k_trsm_reg(const double *Rm, double *X, int N, int rs, int cs, int mode,
                                                  const double *dsc, const int *perm, double *Y) {
  extern __shared__ unsigned char smem_raw[];
  double *rt = reinterpret_cast<double *>(smem_raw);  // [16][N]: columns 16 jk + kk of R
  const int chain = blockIdx.y;
  const double *R = Rm + (size_t)chain * N * N;
  double *Xc = X + (size_t)chain * N * N;
  const int lane = threadIdx.x & 63, l16 = lane & 15;
  const int c = blockIdx.x * 16 + (threadIdx.x >> 4);
  const bool cin = c < N;
  double xr[MJ];
#pragma unroll
  for (int j = 0; j < MJ; ++j) {
    const int i = l16 + 16 * j;
    xr[j] = (cin && i < N) ? Xc[(size_t)i * rs + (size_t)c * cs] : 0.0;
  }
  if (dsc) {
    const double *dc = dsc + (size_t)chain * N;
#pragma unroll
    for (int j = 0; j < MJ; ++j) {
      const int i = l16 + 16 * j;
      if (cin && i < N) xr[j] = xr[j] / dc[i];
    }
  }
  // N <= 256 (one row of R per thread): the next block's 16 columns are read into
  // registers while this block is solved, and stored to LDS at the next step
  constexpr bool PREF = MJ <= 16;
  double vn[16];
  auto read_tile = [&](int jk) {  // all 16 reads in flight (clamped row and column)
    const int i = min((int)threadIdx.x, N - 1);
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) vn[kk] = R[i + (size_t)min(16 * jk + kk, N - 1) * N];
  };
  if constexpr (PREF) read_tile((N - 1) / 16);
  for (int jk = (N - 1) / 16; jk >= 0; --jk) {
    __syncthreads();
    if constexpr (PREF) {
      if ((int)threadIdx.x < N)
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) rt[kk * N + threadIdx.x] = 16 * jk + kk < N ? vn[kk] : 0.0;
    } else {
      for (int i = threadIdx.x; i < N; i += blockDim.x) {
        double v[16];  // all 16 reads in flight before any is used (clamped column, zeroed below)
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) v[kk] = R[i + (size_t)min(16 * jk + kk, N - 1) * N];
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) rt[kk * N + i] = 16 * jk + kk < N ? v[kk] : 0.0;
      }
    }
    __syncthreads();
    if constexpr (PREF) {
      if (jk > 0) read_tile(jk - 1);
    }
    // the block's rows (slot jk) in one register
    double xb = 0.0;
#pragma unroll
    for (int j = 0; j < MJ; ++j)
      if (j == jk) xb = xr[j];
    const int ib = 16 * jk + l16;  // this lane's row in the block
    double pv[16];
#pragma unroll
    for (int kk = 15; kk >= 0; --kk) {
      const int k = 16 * jk + kk;
      if (k >= N) {
        pv[kk] = 0.0;
        continue;
      }
      const double *rk = rt + kk * N;
      const double rkk = rk[k];
      const double xo = mode == 0 ? xb / rkk : (1.0 / rkk) * xb;  // meaningful in lane kk
      if (l16 == kk) xb = xo;
      const double xk = trsm_row_bcast(xo, kk);  // lane kk's solved value to its 16-lane row
      pv[kk] = xk;
      if (l16 < kk) {
        const double r = rk[ib];
        xb = mode == 0 ? xb - xk * r : xb - r * xk;
      }
    }
#pragma unroll
    for (int j = 0; j < MJ; ++j)
      if (j == jk) xr[j] = xb;
    // rows above the block: updates k = 16 jk + 15 .. 16 jk in that order
#pragma unroll
    for (int j = 0; j < MJ; ++j) {
      if (j >= jk) continue;  // uniform
      const int i = l16 + 16 * j;
      double x = xr[j];
#pragma unroll
      for (int kk = 15; kk >= 0; --kk) {
        const double r = rt[kk * N + i];
        x = mode == 0 ? x - pv[kk] * r : x - r * pv[kk];
      }
      xr[j] = x;
    }
  }
  const int *pc = perm ? perm + (size_t)chain * N : nullptr;
  double *Yc = perm ? Y + (size_t)chain * N * N : Xc;
#pragma unroll
  for (int j = 0; j < MJ; ++j) {
    const int i = l16 + 16 * j;
    if (cin && i < N) Yc[(size_t)(pc ? pc[i] - 1 : i) * rs + (size_t)c * cs] = xr[j];
  }
}

// ------------------------------------------------------------------ helpers
// out = op(in) with optional permutation / scaling, one element per thread.
// mode:
//  0 copy                      out[r,c] = in[r,c]
//  1 transpose                 out[r,c] = in[c,r]
//  2 scale columns by d        out[r,c] = in[r,c] * d[c]
//  3 transpose, scale cols     out[r,c] = in[c,r] * d[c]
//  4 rows forward  (ZLAPMR .true.)   out[r,c] = in[p[r]-1, c]
//  5 cols forward  (ZLAPMT .true.)   out[r,c] = in[r, p[c]-1]
//  6 rows backward (ZLAPMR .false.)  out[p[r]-1, c] = in[r, c]
//  7 cols backward (ZLAPMT .false.)  out[r, p[c]-1] = in[r, c]
//  8 divide rows by d          out[r,c] = in[r,c] / d[r]
//  9 divide cols by d          out[r,c] = in[r,c] / d[c]
struct ElemArgs {
  const double *in;
  double *out;
  const double *d;  // chain stride dstride * N
  const int *p;     // chain stride N
  int N, mode;
  int dstride;
};

__global__ void __launch_bounds__(256) k_elem(ElemArgs e) {
  const int N = e.N;
  const int chain = blockIdx.y;
  const double *in = e.in + (size_t)chain * N * N;
  double *out = e.out + (size_t)chain * N * N;
  const double *d = e.d ? e.d + (size_t)chain * e.dstride * N : nullptr;
  const int *p = e.p ? e.p + (size_t)chain * N : nullptr;
  // a block walks whole columns (no index division; coalesced in r)
  for (int c = blockIdx.x; c < N; c += gridDim.x)
  for (int r = threadIdx.x; r < N; r += blockDim.x) {
    const size_t idx = r + (size_t)c * N;
    switch (e.mode) {
      case 0: out[idx] = in[idx]; break;
      case 1: out[idx] = in[c + (size_t)r * N]; break;
      case 2: out[idx] = in[idx] * d[c]; break;
      case 3: out[idx] = in[c + (size_t)r * N] * d[c]; break;
      case 4: out[idx] = in[(p[r] - 1) + (size_t)c * N]; break;
      case 5: out[idx] = in[r + (size_t)(p[c] - 1) * N]; break;
      case 6: out[(p[r] - 1) + (size_t)c * N] = in[idx]; break;
      case 7: out[r + (size_t)(p[c] - 1) * N] = in[idx]; break;
      case 8: out[idx] = in[idx] / d[r]; break;
      case 9: out[idx] = in[idx] / d[c]; break;
    }
  }
}

// M = X + (DR VRVL DL), transposed if trans (src/stabilization.f90:138-148)
__global__ void __launch_bounds__(256) k_build_m(const double *X, const double *VRVL, const double *DR,
                                                 const double *DL, double *M, int N, int trans) {
  const int chain = blockIdx.y;
  const double *x = X + (size_t)chain * N * N, *v = VRVL + (size_t)chain * N * N;
  const double *dr = DR + (size_t)chain * N, *dl = DL + (size_t)chain * N;
  double *m = M + (size_t)chain * N * N;
  for (int c = blockIdx.x; c < N; c += gridDim.x)
  for (int r = threadIdx.x; r < N; r += blockDim.x) {
    const size_t idx = r + (size_t)c * N;
    double t = 0.0 + dl[c] * v[idx];
    double w = 0.0 + dr[r] * t;
    double val = x[idx] + w;
    if (trans) m[c + (size_t)r * N] = val;
    else m[idx] = val;
  }
}

// D-split scales (Loh et al.): for D > 0, Db = max(D, 1), Ds = min(D, 1), so D = Db Ds.
// out per chain: [0,N) DRb, [N,2N) DRs, [2N,3N) DLb, [3N,4N) DLs.
__global__ void __launch_bounds__(256) k_dsplit(const double *DR, const double *DL, double *out, int N) {
  const int chain = blockIdx.x;
  const double *dr = DR + (size_t)chain * N, *dl = DL + (size_t)chain * N;
  double *o = out + (size_t)chain * 4 * N;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    o[i] = dr[i] > 1.0 ? dr[i] : 1.0;
    o[N + i] = dr[i] > 1.0 ? 1.0 : dr[i];
    o[2 * N + i] = dl[i] > 1.0 ? dl[i] : 1.0;
    o[3 * N + i] = dl[i] > 1.0 ? 1.0 : dl[i];
  }
}

// M' = DRb^{-1} X DLb^{-1} + DRs VRVL DLs (the D-split form of the matrix of
// stab_green: all entries bounded by the scales' ratio 1, no huge D products)
__global__ void __launch_bounds__(256) k_build_m_split(const double *X, const double *VRVL, const double *sc,
                                                       double *M, int N) {
  const int chain = blockIdx.y;
  const double *x = X + (size_t)chain * N * N, *v = VRVL + (size_t)chain * N * N;
  const double *s = sc + (size_t)chain * 4 * N;
  double *m = M + (size_t)chain * N * N;
  for (int c = blockIdx.x; c < N; c += gridDim.x)
    for (int r = threadIdx.x; r < N; r += blockDim.x) {
      const size_t idx = r + (size_t)c * N;
      const double a = (x[idx] / s[r]) / s[2 * N + c];
      const double b = s[N + r] * (v[idx] * s[3 * N + c]);
      m[idx] = a + b;
    }
}

// drift = max |Gnew - Gold| per chain; Xmaxm/Xmeanm update; Gold <- Gnew.
// gridDim.x workgroups per chain, each over a contiguous part of the matrix; the
// partial maxima go to part[chain * pstride + b] and k_drift_finish takes their
// maximum (exact, so the split does not change the result).
template <typename R>
__global__ void __launch_bounds__(256) k_drift_replace(const double *Gnew, R *Gold, double *part, int pstride,
                                                       int N) {
  const int chain = blockIdx.y;
  const size_t nn = (size_t)N * N, per = (nn + gridDim.x - 1) / gridDim.x;
  const size_t lo = (size_t)blockIdx.x * per, hi = lo + per < nn ? lo + per : nn;
  const double *gn = Gnew + (size_t)chain * nn;
  R *go = Gold + (size_t)chain * nn;
  double m = 0.0;
  for (size_t idx = lo + threadIdx.x; idx < hi; idx += blockDim.x) {
    double d = fabs(gn[idx] - (double)go[idx]);
    m = d > m ? d : m;
    go[idx] = (R)gn[idx];
  }
  for (int off = 32; off > 0; off >>= 1) {
    double o = __shfl_xor(m, off, 64);
    m = o > m ? o : m;
  }
  __shared__ double red[16];
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double mm = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) mm = red[w] > mm ? red[w] : mm;
    part[(size_t)chain * pstride + blockIdx.x] = mm;
  }
}

__global__ void __launch_bounds__(64) k_drift_finish(const double *part, int pstride, int nb, double *xmax,
                                                     double *xmean, int nch, int accumulate) {
  const int chain = blockIdx.x * blockDim.x + threadIdx.x;
  if (chain >= nch) return;
  double mm = 0.0;
  for (int b = 0; b < nb; ++b) {
    const double v = part[(size_t)chain * pstride + b];
    mm = v > mm ? v : mm;
  }
  if (mm > xmax[chain]) xmax[chain] = mm;
  if (accumulate) xmean[chain] += mm;
}
