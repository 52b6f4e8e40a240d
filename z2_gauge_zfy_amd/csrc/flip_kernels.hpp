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
// the serial one.  The recorded updates are applied to G by this kernel as one
// rank-2k product on the matrix cores whenever k = `kdelay` of them are pending
// (the delay depth), and the remaining ones by the next kinetic kernel (flush
// prologue, in acceptance order).
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
  int final_flush;        // FP32 only, 1: flush the updates still pending at the end into G and
                          // write (g*ep)*em (right pass, LocalK_prop_R then propT_R); 2: flush
                          // without the mu factors (LocalK_prop_R alone); nacc is then 0
  const uint32_t *force;  // optional: decisions to replay, bit k = proposal k, chain stride dwords
  uint32_t *record;       // optional: decisions taken, same layout
  int dwords;
  LatticeDev lat;
  // Wolff trial pass (ratioK_fermion, src/globalK.f90:15-109): when gbond is set
  // the flip of bond gbond[chain] is applied at the slices where gmask is 1 and
  // every other proposal is rejected; no random number is drawn, no counter
  // moves, and log max(|det|^2, 1e-300) of each applied flip is added to glr[2*chain]
  const int *gbond;       // per chain, 0-based
  const int *gmask;       // per chain T
  double *glr;            // per chain [2]
};

template <typename R>
__device__ __forceinline__ R flip_read(const R *G, int i, int j, int N, int mu, double ep, double em) {
  R x = G[i + (size_t)j * N];
  if (mu) {
    x = x * (R)ep;
    x = x * (R)em;
  }
  return x;
}

// LDS bytes of k_flip for N sites and delay depth kd (the host sizes the launch with it)
__host__ __device__ inline size_t flip_lds_bytes(int N, int kd) {
  const size_t N2 = 2 * (size_t)N;
  size_t b = sizeof(double) * (2 * (size_t)N + 3 * N2 + 4 * (size_t)N * kd + 16);
  b += sizeof(short) * 8 * N2;  // bond sites (2) and plaquette partner bonds (6) per bond
  b += 2 * N2;                  // current field, sum of the neighbouring slices' fields
  return (b + 15) & ~(size_t)15;
}

// 16x16x4 matrix-core product in the precision of the propagated G:
// lane l supplies A[l%16][l/16] and B[l/16][l%16]; item q holds
// D[mfma_row(q, l/16)][l%16]
template <typename R> struct MfmaAcc;
template <> struct MfmaAcc<double> { typedef double type __attribute__((ext_vector_type(4))); };
template <> struct MfmaAcc<float> { typedef float type __attribute__((ext_vector_type(4))); };
// output row of item q in lane group lk = lane/16 (the two precisions differ)
template <typename R> __device__ __forceinline__ int mfma_row(int q, int lk);
template <> __device__ __forceinline__ int mfma_row<double>(int q, int lk) { return 4 * q + lk; }
template <> __device__ __forceinline__ int mfma_row<float>(int q, int lk) { return 4 * lk + q; }
#define FLIP_BPW 4  // flush: column blocks per wavefront task
template <typename R>
__device__ __forceinline__ typename MfmaAcc<R>::type mfma_16x16x4(R a, R b, typename MfmaAcc<R>::type c);
template <>
__device__ __forceinline__ MfmaAcc<double>::type mfma_16x16x4<double>(double a, double b, MfmaAcc<double>::type c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
template <>
__device__ __forceinline__ MfmaAcc<float>::type mfma_16x16x4<float>(float a, float b, MfmaAcc<float>::type c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// minimum over the 64 lanes of a wavefront (all lanes active), in every lane
__device__ __forceinline__ int wave_min_int(int v) {
  v = min(v, __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false));   // quad_perm(1,0,3,2)
  v = min(v, __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false));   // quad_perm(2,3,0,1)
  v = min(v, __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, false));  // row_half_mirror
  v = min(v, __builtin_amdgcn_mov_dpp(v, 0x140, 0xF, 0xF, false));  // row_mirror
  return min(min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
             min(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

// MPT: proposals per thread (2N <= 512 MPT); a thread keeps the lattice data,
// random number and current 2x2 block of its proposals in registers.
// position of element i of vector c (0, 1) of a pending update's T (V: after 2N):
// interleaved (2 i + c) in FP64, where the update vectors read the two together,
// blocked (c N + i) in FP32, where the flush's operand reads stay unit-stride
template <typename R>
__device__ __forceinline__ int pt_ix(int N, int c, int i) {
  return sizeof(R) == 8 ? 2 * i + c : c * N + i;
}

// NC: the site count when fixed at compile time (0: a.lat.N), so that the
// distances between the pending vectors and G's columns are constants
template <typename R, int MPT, int NC = 0>
__global__ void __launch_bounds__(512) k_flip(FlipArgs<R> a) {
  const int N = NC ? NC : a.lat.N, N2 = NC ? 2 * NC : a.lat.N2;
  const int chain = blockIdx.x;
  const int tid = threadIdx.x, nth = blockDim.x;
  const ChainConst *cp = a.cc + chain;
  const double ep = cp->ep, em = cp->em;
  R *G = a.G + (size_t)chain * N * N;
  int8_t *sigc = a.sig + (size_t)chain * a.T * N2;
  int8_t *sig_t = sigc + (size_t)(a.nt - 1) * N2;
  const int tp = (a.nt - 2 + a.T) % a.T, tn = a.nt % a.T;  // 0-based neighbour slices
  const int8_t *sig_p = sigc + (size_t)tp * N2, *sig_n = sigc + (size_t)tn * N2;
  R *Tb = a.Tbuf + chain * a.upd_stride;
  R *Vb = a.Vbuf + chain * a.upd_stride;

  extern __shared__ unsigned char smem_raw[];
  double *dgb = reinterpret_cast<double *>(smem_raw);  // 2 x N : G_ii (double-buffered per round)
  double *o12 = dgb + 2 * N;                           // 2N    : G_p1p2 per bond
  double *o21 = o12 + N2;                              // 2N    : G_p2p1 per bond
  double *rnd = o21 + N2;                              // 2N    : random number per proposal index
  double *PT = rnd + N2;                               // kdelay x 4N : pending updates, slot k:
                                                       // T_c(i) at 4Nk + 2i + c, V_c(i) at 4Nk + 2N + 2i + c
  double *bos = PT + (size_t)4 * N * a.kdelay;         // 9 (16) : boson ratio table
  short *bp = reinterpret_cast<short *>(bos + 16);     // 2N x 2 : bond sites
  short *bbs = bp + 2 * N2;                            // 2N x 6 : plaquette partner bonds
  int8_t *sg = reinterpret_cast<int8_t *>(bbs + 6 * N2);  // 2N : current slice field
  int8_t *sgs = sg + N2;                                  // 2N : sig(nt-1) + sig(nt+1)
  __shared__ int s_kmin[2];
  __shared__ uint32_t s_dec[64];  // decision bits of this slice (2N <= 2048)

  // delta for s_old = +1 / -1 in registers (no dynamic indexing of the constants)
  const double Dp0 = cp->delta[0][0], Dp1 = cp->delta[0][1], Dp2 = cp->delta[0][2], Dp3 = cp->delta[0][3];
  const double Dm0 = cp->delta[1][0], Dm1 = cp->delta[1][1], Dm2 = cp->delta[1][2], Dm3 = cp->delta[1][3];

  int mu = a.mu_on_read;
  int mu_done = 0;
  for (int i = tid; i < N; i += nth) dgb[i] = (double)flip_read<R>(G, i, i, N, mu, ep, em);
  const uint32_t seed0 = (uint32_t)a.seed[chain] & 0x7fffffffu;
  for (int b = tid; b < N2; b += nth) {
    const int p1 = a.lat.bond_p[2 * b], p2 = a.lat.bond_p[2 * b + 1];
    bp[2 * b] = (short)p1;
    bp[2 * b + 1] = (short)p2;
    for (int q = 0; q < 6; ++q) bbs[6 * b + q] = (short)a.lat.bond_bonds[6 * b + q];
    o12[b] = (double)flip_read<R>(G, p1, p2, N, mu, ep, em);
    o21[b] = (double)flip_read<R>(G, p2, p1, N, mu, ep, em);
    sg[b] = sig_t[b];
    sgs[b] = (int8_t)((int)sig_p[b] + (int)sig_n[b]);
    // proposal index k = b: draw number k+1 of this slice
    uint64_t x = (uint64_t)seed0 * (uint64_t)a.apow[b + 1];
    rnd[b] = (double)(uint32_t)(x & 0x7fffffffull) / 2147483648.0;
  }
  if (tid < 9) bos[tid] = cp->boson[tid];
  for (int q = tid; q < 64; q += nth) s_dec[q] = 0u;
  if (tid == 0) {
    s_kmin[0] = 2 * N2 + 2;
    s_kmin[1] = 2 * N2 + 2;
  }
  const uint32_t *frc = a.force ? a.force + (size_t)chain * a.dwords : nullptr;
  // forced bond of a Wolff trial at this slice (-1: none)
  const int gb = a.gbond ? (a.gmask[(size_t)chain * a.T + (a.nt - 1)] ? a.gbond[chain] : -1) : -1;
  __syncthreads();
  // per-thread proposals k = tid + m * nth: bond, sites, partner bonds, field,
  // random number, 2x2 block (G_p1p1, G_p2p2, G_p1p2, G_p2p1), all in registers
  int rb_[MPT], rp1[MPT], rp2[MPT], rbb[MPT][6], rso[MPT], rss[MPT];
  double rrn[MPT], r11[MPT], r22[MPT], r12[MPT], r21[MPT];
#pragma unroll
  for (int m = 0; m < MPT; ++m) {
    const int k = tid + m * nth;
    const int kc = k < N2 ? k : 0;
    const int b = a.descending ? (N2 - 1 - kc) : kc;
    rb_[m] = b;
    rp1[m] = bp[2 * b];
    rp2[m] = bp[2 * b + 1];
#pragma unroll
    for (int q = 0; q < 6; ++q) rbb[m][q] = bbs[6 * b + q];
    rso[m] = sg[b];
    rss[m] = sgs[b];
    rrn[m] = rnd[kc];
    r11[m] = dgb[rp1[m]];
    r22[m] = dgb[rp2[m]];
    r12[m] = o12[b];
    r21[m] = o21[b];
  }

  // Rounds: (A) flush if kdelay updates are pending, fold the last accepted
  // update into the cached 2x2 blocks and evaluate the remaining proposals;
  // (B) build the rank-2 update of the first accepted one.  The values a
  // proposal sees are bit-identical to those of the serial loop.
  int start = 0, na = 0, last = -1, cur = 0, round = 0, tot = 0;
  // FP32 flush of the na pending updates into G (mu_w: write (g*ep)*em), used in
  // the round loop and, for the right pass, once more at the end; the FP64 flush
  // stays inline in the round loop
  auto flush32 = [&](const bool mu_w) {
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63, nw = nth >> 6;
    const int li = lane & 15, lk = lane >> 4;
    constexpr int BPW = FLIP_BPW;
    const int nbk = (N + 15) >> 4, ncg = (nbk + BPW - 1) / BPW, K = 2 * na;
    // FP32: O = -T V^T, item q of lane (li, lk) is G(16 rb + 4 lk + q, col):
    // four consecutive rows of one column, one 16-byte access per tile
    // the tiles of the wavefront's next task are requested before this task's
    // products and stores, so waiting for them never waits for the stores
    if (MPT < 4 && K <= 32 && ncg <= nw && nw % ncg == 0) {  // (registers: not with 4 proposals per thread)
      // a wavefront keeps one group of BPW column blocks: its -V operands (K <= 32:
      // eight K steps) are read from LDS once into registers, then its row blocks
      // are swept; every output block sees the same products in the same order
      const int wpg = nw / ncg, g = wave % ncg, rb0 = wave / ncg;
      const int cb0 = BPW * 16 * g;
      float bvr[8][BPW];
#pragma unroll
      for (int st = 0; st < 8; ++st) {
        const int kq = 4 * st + lk, kc = kq < K ? kq : 0;
        const double *P = PT + (size_t)4 * N * (kc >> 1);
        const int pc = kc & 1;
#pragma unroll
        for (int u = 0; u < BPW; ++u) {
          const int c = cb0 + 16 * u + li;
          const float v = (float)(-P[2 * N + pt_ix<R>(N, pc, c < N ? c : 0)]);
          bvr[st][u] = (kq < K && c < N) ? v : 0.f;  // -V_kq(c)
        }
      }
      auto load_rb = [&](int rb, float4 *tile) {
        const int row0 = 16 * rb + 4 * lk;
#pragma unroll
        for (int u = 0; u < BPW; ++u) {
          const int col = cb0 + 16 * u + li;
          const bool in = row0 < N && col < N;
          tile[u] = *reinterpret_cast<const float4 *>(G + (in ? row0 + (size_t)col * N : 0));  // selected at use
        }
      };
      float4 cur[BPW], nxt[BPW];
      if (rb0 < nbk) load_rb(rb0, cur);
      for (int rb = rb0; rb < nbk; rb += wpg) {
        if (rb + wpg < nbk) load_rb(rb + wpg, nxt);
        const int row0 = 16 * rb + 4 * lk;
        typename MfmaAcc<R>::type acc[BPW];
#pragma unroll
        for (int u = 0; u < BPW; ++u) {
          const bool in = row0 < N && cb0 + 16 * u + li < N;
          float x0 = in ? cur[u].x : 0.f, x1 = in ? cur[u].y : 0.f, x2 = in ? cur[u].z : 0.f, x3 = in ? cur[u].w : 0.f;
          if (mu) {
            x0 = x0 * (float)ep;
            x0 = x0 * (float)em;
            x1 = x1 * (float)ep;
            x1 = x1 * (float)em;
            x2 = x2 * (float)ep;
            x2 = x2 * (float)em;
            x3 = x3 * (float)ep;
            x3 = x3 * (float)em;
          }
          acc[u][0] = x0;
          acc[u][1] = x1;
          acc[u][2] = x2;
          acc[u][3] = x3;
        }
        const int r = 16 * rb + li;
#pragma unroll
        for (int st = 0; st < 8; ++st) {
          if (4 * st < K) {
            const int kq = 4 * st + lk, kc = kq < K ? kq : 0;
            const double *P = PT + (size_t)4 * N * (kc >> 1);
            const int pc = kc & 1;
            const float t = (float)P[pt_ix<R>(N, pc, r < N ? r : 0)];
            const R av = (kq < K && r < N) ? t : (R)0;  // T_kq(r)
#pragma unroll
            for (int u = 0; u < BPW; ++u) acc[u] = mfma_16x16x4<R>(av, bvr[st][u], acc[u]);
          }
        }
#pragma unroll
        for (int u = 0; u < BPW; ++u) {
          const int col = cb0 + 16 * u + li;
          if (mu_w)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              acc[u][q] = acc[u][q] * (float)ep;
              acc[u][q] = acc[u][q] * (float)em;
            }
          if (row0 < N && col < N)
            *reinterpret_cast<float4 *>(G + row0 + (size_t)col * N) =
                make_float4(acc[u][0], acc[u][1], acc[u][2], acc[u][3]);
        }
#pragma unroll
        for (int u = 0; u < BPW; ++u) cur[u] = nxt[u];
      }
      return;
    }
    const int ntask = nbk * ncg;
    auto load_tiles = [&](int task, float4 *tile) {
      const int row0 = 16 * (task % nbk) + 4 * lk;  // N is a multiple of 4: a float4 is all in or all out
      const int cb0 = BPW * 16 * (task / nbk);
#pragma unroll
      for (int u = 0; u < BPW; ++u) {
        const int col = cb0 + 16 * u + li;
        const bool in = row0 < N && col < N;
        tile[u] = *reinterpret_cast<const float4 *>(G + (in ? row0 + (size_t)col * N : 0));  // selected at use
      }
    };
    float4 cur[BPW], nxt[BPW];
    if (wave < ntask) load_tiles(wave, cur);
    for (int task = wave; task < ntask; task += nw) {
      if (task + nw < ntask) load_tiles(task + nw, nxt);
      const int rb = task % nbk;
      const int cb0 = BPW * 16 * (task / nbk);
      const int row0 = 16 * rb + 4 * lk;
      typename MfmaAcc<R>::type acc[BPW];
#pragma unroll
      for (int u = 0; u < BPW; ++u) {
        const bool in = row0 < N && cb0 + 16 * u + li < N;
        float x0 = in ? cur[u].x : 0.f, x1 = in ? cur[u].y : 0.f, x2 = in ? cur[u].z : 0.f, x3 = in ? cur[u].w : 0.f;
        if (mu) {
          x0 = x0 * (float)ep;
          x0 = x0 * (float)em;
          x1 = x1 * (float)ep;
          x1 = x1 * (float)em;
          x2 = x2 * (float)ep;
          x2 = x2 * (float)em;
          x3 = x3 * (float)ep;
          x3 = x3 * (float)em;
        }
        acc[u][0] = x0;
        acc[u][1] = x1;
        acc[u][2] = x2;
        acc[u][3] = x3;
      }
      const int r = 16 * rb + li;
      for (int kk = 0; kk < K; kk += 4) {
        const int kq = kk + lk;
        const double *P = PT + (size_t)4 * N * (kq >> 1);
        const int pc = kq & 1;
        const R av = (kq < K && r < N) ? (R)P[pt_ix<R>(N, pc, r)] : (R)0;  // T_kq(r)
#pragma unroll
        for (int u = 0; u < BPW; ++u) {
          const int c = cb0 + 16 * u + li;
          const R bv = (kq < K && c < N) ? (R)(-P[2 * N + pt_ix<R>(N, pc, c)]) : (R)0;  // -V_kq(c)
          acc[u] = mfma_16x16x4<R>(av, bv, acc[u]);
        }
      }
#pragma unroll
      for (int u = 0; u < BPW; ++u) {
        const int col = cb0 + 16 * u + li;
        if (mu_w)
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            acc[u][q] = acc[u][q] * (float)ep;
            acc[u][q] = acc[u][q] * (float)em;
          }
        if (row0 < N && col < N)
          *reinterpret_cast<float4 *>(G + row0 + (size_t)col * N) =
              make_float4(acc[u][0], acc[u][1], acc[u][2], acc[u][3]);
      }
#pragma unroll
      for (int u = 0; u < BPW; ++u) cur[u] = nxt[u];
    }
  };
  while (true) {
    const int slot = round & 1;
    if (na >= a.kdelay) {
      // flush the pending updates into G as one rank-2na product on the matrix
      // cores (and mu, once): G(r, c) -= sum_k T_k(r) V_k(c).  Output blocks are
      // computed transposed (O = -V T^T) so that the 16 lanes of an output row
      // cover 16 consecutive rows of G; a wavefront keeps FLIP_BPW column blocks
      // of one row block in flight (they share the T operand).
      const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63, nw = nth >> 6;
      const int li = lane & 15, lk = lane >> 4;
      constexpr int BPW = FLIP_BPW;
      const int nbk = (N + 15) >> 4, ncg = (nbk + BPW - 1) / BPW, K = 2 * na;
      if constexpr (sizeof(R) == 4) {
        flush32(false);
      } else {
        bool done = false;
        if constexpr (MPT == 1) {
          if (K <= 32 && ncg <= nw && nw % ncg == 0) {
            // as in the FP32 flush: a wavefront keeps one column group and its -V
            // operands in registers for the whole flush (same products, same order)
            const int wpg = nw / ncg, g = wave % ncg;
            const int cb0 = BPW * 16 * g;
            double avr[8][BPW];
  #pragma unroll
            for (int st = 0; st < 8; ++st) {
              const int kq = 4 * st + lk, kc = kq < K ? kq : 0;
              const double *P = PT + (size_t)4 * N * (kc >> 1);
              const int pc = kc & 1;
  #pragma unroll
              for (int u = 0; u < BPW; ++u) {
                const int c = cb0 + 16 * u + li;
                const double v = -P[2 * N + pt_ix<R>(N, pc, c < N ? c : 0)];
                avr[st][u] = (kq < K && c < N) ? v : 0.0;  // -V_kq(c)
              }
            }
            for (int rb = wave / ncg; rb < nbk; rb += wpg) {
              const int r = 16 * rb + li;
              typename MfmaAcc<R>::type acc[BPW];
  #pragma unroll
              for (int u = 0; u < BPW; ++u)
  #pragma unroll
                for (int q = 0; q < 4; ++q) {
                  const int col = cb0 + 16 * u + mfma_row<R>(q, lk);
                  R x = (r < N && col < N) ? G[r + (size_t)col * N] : (R)0;
                  if (mu) {
                    x = x * (R)ep;
                    x = x * (R)em;
                  }
                  acc[u][q] = x;
                }
  #pragma unroll
              for (int st = 0; st < 8; ++st) {
                if (4 * st < K) {
                  const int kq = 4 * st + lk, kc = kq < K ? kq : 0;
                  const double *P = PT + (size_t)4 * N * (kc >> 1);
                  const int pc = kc & 1;
                  const double t = P[pt_ix<R>(N, pc, r < N ? r : 0)];
                  const R bv = (kq < K && r < N) ? (R)t : (R)0;  // T_kq(r)
  #pragma unroll
                  for (int u = 0; u < BPW; ++u) acc[u] = mfma_16x16x4<R>((R)avr[st][u], bv, acc[u]);
                }
              }
  #pragma unroll
              for (int u = 0; u < BPW; ++u)
  #pragma unroll
                for (int q = 0; q < 4; ++q) {
                  const int col = cb0 + 16 * u + mfma_row<R>(q, lk);
                  if (r < N && col < N) G[r + (size_t)col * N] = acc[u][q];
                }
            }
            done = true;
          }
        }
        if (!done)
        for (int task = wave; task < nbk * ncg; task += nw) {
          const int r = 16 * (task % nbk) + li;
          const int cb0 = BPW * 16 * (task / nbk);
          typename MfmaAcc<R>::type acc[BPW];
  #pragma unroll
          for (int u = 0; u < BPW; ++u)
  #pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int col = cb0 + 16 * u + mfma_row<R>(q, lk);
              R x = (r < N && col < N) ? G[r + (size_t)col * N] : (R)0;
              if (mu) {
                x = x * (R)ep;
                x = x * (R)em;
              }
              acc[u][q] = x;
            }
          for (int kk = 0; kk < K; kk += 4) {
            const int kq = kk + lk;
            const double *P = PT + (size_t)4 * N * (kq >> 1);
            const int pc = kq & 1;
            const R bv = (kq < K && r < N) ? (R)P[pt_ix<R>(N, pc, r)] : (R)0;  // T_kq(r)
  #pragma unroll
            for (int u = 0; u < BPW; ++u) {
              const int c = cb0 + 16 * u + li;
              const R av = (kq < K && c < N) ? (R)(-P[2 * N + pt_ix<R>(N, pc, c)]) : (R)0;  // -V_kq(c)
              acc[u] = mfma_16x16x4<R>(av, bv, acc[u]);
            }
          }
  #pragma unroll
          for (int u = 0; u < BPW; ++u)
  #pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int col = cb0 + 16 * u + mfma_row<R>(q, lk);
              if (r < N && col < N) G[r + (size_t)col * N] = acc[u][q];
            }
        }
      }
      if (mu) mu_done = 1;
      mu = 0;
      na = 0;
    }
    const double *dgc = dgb + (size_t)cur * N;
    double *dgn = dgb + (size_t)(cur ^ 1) * N;
    const double *TA = last >= 0 ? PT + (size_t)4 * N * last : nullptr;
    const double *VA = last >= 0 ? TA + 2 * N : nullptr;
    if (last >= 0)
      for (int i = tid; i < N; i += nth) {
        double d = TA[pt_ix<R>(N, 0, i)] * VA[pt_ix<R>(N, 0, i)] + TA[pt_ix<R>(N, 1, i)] * VA[pt_ix<R>(N, 1, i)];
        dgn[i] = dgc[i] - d;
      }
    int cand = 2 * N2 + 2;
#pragma unroll
    for (int m = 0; m < MPT; ++m) {
      const int k = tid + m * nth;
      if (k >= N2 || k < start) continue;
      const int p1 = rp1[m], p2 = rp2[m];
      if (last >= 0) {
        double d12 = TA[pt_ix<R>(N, 0, p1)] * VA[pt_ix<R>(N, 0, p2)] + TA[pt_ix<R>(N, 1, p1)] * VA[pt_ix<R>(N, 1, p2)];
        double d21 = TA[pt_ix<R>(N, 0, p2)] * VA[pt_ix<R>(N, 0, p1)] + TA[pt_ix<R>(N, 1, p2)] * VA[pt_ix<R>(N, 1, p1)];
        r12[m] = r12[m] - d12;
        r21[m] = r21[m] - d21;
        o12[rb_[m]] = r12[m];
        o21[rb_[m]] = r21[m];
        double e1 = TA[pt_ix<R>(N, 0, p1)] * VA[pt_ix<R>(N, 0, p1)] + TA[pt_ix<R>(N, 1, p1)] * VA[pt_ix<R>(N, 1, p1)];
        double e2 = TA[pt_ix<R>(N, 0, p2)] * VA[pt_ix<R>(N, 0, p2)] + TA[pt_ix<R>(N, 1, p2)] * VA[pt_ix<R>(N, 1, p2)];
        r11[m] = r11[m] - e1;
        r22[m] = r22[m] - e2;
      }
      if (cand < 2 * N2) continue;  // an earlier proposal of this thread is accepted
      const int s_old = rso[m];
      const double D0 = s_old > 0 ? Dp0 : Dm0, D1 = s_old > 0 ? Dp1 : Dm1;
      const double D2 = s_old > 0 ? Dp2 : Dm2, D3 = s_old > 0 ? Dp3 : Dm3;
      const double A11 = 1.0 - r11[m], A21 = 0.0 - r21[m];
      const double A12 = 0.0 - r12[m], A22 = 1.0 - r22[m];
      const double P11 = 1.0 + (D0 * A11 + D2 * A21);
      const double P21 = 0.0 + (D1 * A11 + D3 * A21);
      const double P12 = 0.0 + (D0 * A12 + D2 * A22);
      const double P22 = 1.0 + (D1 * A12 + D3 * A22);
      const double det = P11 * P22 - P12 * P21;
      // boson ratio: plaquette sums (src/fields.f90:66-101) and temporal term (:104-117)
      const int s_new = -s_old;
      int tn1 = s_new * (int)sg[rbb[m][0]] * (int)sg[rbb[m][1]] * (int)sg[rbb[m][2]];
      int tn2 = s_new * (int)sg[rbb[m][3]] * (int)sg[rbb[m][4]] * (int)sg[rbb[m][5]];
      int dS = 2 * (tn1 + tn2);  // tmp_new - tmp_old with tmp_old = -tmp_new
      int dT = 2 * s_new * rss[m];
      const double rb = bos[(dS / 4 + 1) * 3 + (dT / 4 + 1)];
      const double ratio = (det * det) * rb;
      const bool acc = a.gbond ? rb_[m] == gb
                               : (frc ? ((frc[k >> 5] >> (k & 31)) & 1u) != 0u : fabs(ratio) > rrn[m]);
      if (acc) cand = 2 * k + (s_old > 0 ? 1 : 0);
    }
    cand = wave_min_int(cand);
    if ((tid & 63) == 0 && cand < 2 * N2) atomicMin(&s_kmin[slot], cand);
    if (tid == 0) s_kmin[slot ^ 1] = 2 * N2 + 2;  // read by all threads before the last barrier
    __syncthreads();
    if (last >= 0) cur ^= 1;
    const int packed = s_kmin[slot];
    if (packed >= 2 * N2) break;
    const int kmin = packed >> 1, s_old = (packed & 1) ? 1 : -1;
    const int b = a.descending ? (N2 - 1 - kmin) : kmin;
    const int p1 = bp[2 * b], p2 = bp[2 * b + 1];
    // this thread's first G loads of the update, issued before the 2x2 inverse
    R g1 = (R)0, g2 = (R)0;
    if (tid < 2 * N) {
      if (tid < N) {
        g1 = G[tid + (size_t)p1 * N];
        g2 = G[tid + (size_t)p2 * N];
      } else {
        g1 = G[p1 + (size_t)(tid - N) * N];
        g2 = G[p2 + (size_t)(tid - N) * N];
      }
    }
    const double *dgr = dgb + (size_t)cur * N;
    const double D11 = s_old > 0 ? Dp0 : Dm0, D21 = s_old > 0 ? Dp1 : Dm1;
    const double D12 = s_old > 0 ? Dp2 : Dm2, D22 = s_old > 0 ? Dp3 : Dm3;
    double Pi11, Pi21, Pi12, Pi22;
    {
      const double A11 = 1.0 - dgr[p1], A21 = 0.0 - o21[b];
      const double A12 = 0.0 - o12[b], A22 = 1.0 - dgr[p2];
      const double P11 = 1.0 + (D11 * A11 + D12 * A21);
      const double P21 = 0.0 + (D21 * A11 + D22 * A21);
      const double P12 = 0.0 + (D11 * A12 + D12 * A22);
      const double P22 = 1.0 + (D21 * A12 + D22 * A22);
      const double det = P11 * P22 - P12 * P21;
      Pi11 = P22 / det;
      Pi21 = (-P21) / det;
      Pi12 = (-P12) / det;
      Pi22 = P11 / det;
      if (a.gbond && tid == 0) {
        double r = fabs(det * det);
        if (r < 1.0e-300) r = 1.0e-300;
        a.glr[2 * chain] = a.glr[2 * chain] + log(r > 1.0e-300 ? r : 1.0e-300);
      }
    }
    if (tid == 0) {
      sg[b] = (int8_t)(-s_old);
      s_dec[kmin >> 5] |= 1u << (kmin & 31);
      tot = tot + 1;
    }
    double *TN = PT + (size_t)4 * N * na;  // slot of this accept: T0/T1 and V0/V1 interleaved
    for (int q = tid; q < 2 * N; q += nth) {
      if (q < N) {
        // column part: current G(i,p1), G(i,p2) with pending accepts applied in order
        const int i = q;
        R x1 = q == tid ? g1 : G[i + (size_t)p1 * N];
        R x2 = q == tid ? g2 : G[i + (size_t)p2 * N];
        if (mu) {
          x1 = x1 * (R)ep;
          x1 = x1 * (R)em;
          x2 = x2 * (R)ep;
          x2 = x2 * (R)em;
        }
        double c1 = (double)x1;
        double c2 = (double)x2;
        for (int k = 0; k < na; ++k) {
          const double *P = PT + (size_t)4 * N * k;
          double ti0 = P[pt_ix<R>(N, 0, i)], ti1 = P[pt_ix<R>(N, 1, i)];
          double vp10 = P[2 * N + pt_ix<R>(N, 0, p1)], vp11 = P[2 * N + pt_ix<R>(N, 1, p1)], vp20 = P[2 * N + pt_ix<R>(N, 0, p2)], vp21 = P[2 * N + pt_ix<R>(N, 1, p2)];
          c1 = c1 - (ti0 * vp10 + ti1 * vp11);
          c2 = c2 - (ti0 * vp20 + ti1 * vp21);
        }
        TN[pt_ix<R>(N, 0, i)] = c1 * Pi11 + c2 * Pi21;
        TN[pt_ix<R>(N, 1, i)] = c1 * Pi12 + c2 * Pi22;
      } else {
        // row part: current G(p1,i), G(p2,i)
        const int i = q - N;
        R x1 = q == tid ? g1 : G[p1 + (size_t)i * N];
        R x2 = q == tid ? g2 : G[p2 + (size_t)i * N];
        if (mu) {
          x1 = x1 * (R)ep;
          x1 = x1 * (R)em;
          x2 = x2 * (R)ep;
          x2 = x2 * (R)em;
        }
        double r1 = (double)x1;
        double r2 = (double)x2;
        for (int k = 0; k < na; ++k) {
          const double *P = PT + (size_t)4 * N * k;
          double tp10 = P[pt_ix<R>(N, 0, p1)], tp11 = P[pt_ix<R>(N, 1, p1)], tp20 = P[pt_ix<R>(N, 0, p2)], tp21 = P[pt_ix<R>(N, 1, p2)];
          double vi0 = P[2 * N + pt_ix<R>(N, 0, i)], vi1 = P[2 * N + pt_ix<R>(N, 1, i)];
          r1 = r1 - (tp10 * vi0 + tp11 * vi1);
          r2 = r2 - (tp20 * vi0 + tp21 * vi1);
        }
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
        TN[2 * N + pt_ix<R>(N, 0, i)] = v0;
        TN[2 * N + pt_ix<R>(N, 1, i)] = v1;
      }
    }
    __syncthreads();
    last = na;
    na = na + 1;
    start = kmin + 1;
    round = round + 1;
  }
  if constexpr (sizeof(R) == 4) {
    if (a.final_flush) {
      flush32(a.final_flush == 1);
      na = 0;
    }
  }
  // the pending updates left for the next kinetic kernel
  for (int idx = tid; idx < na * N; idx += nth) {
    const int k = idx / N, i = idx - k * N;
    const double *P = PT + (size_t)4 * N * k;
    Tb[(2 * k) * N + i] = (R)P[pt_ix<R>(N, 0, i)];
    Tb[(2 * k + 1) * N + i] = (R)P[pt_ix<R>(N, 1, i)];
    Vb[(2 * k) * N + i] = (R)P[2 * N + pt_ix<R>(N, 0, i)];
    Vb[(2 * k + 1) * N + i] = (R)P[2 * N + pt_ix<R>(N, 1, i)];
  }
  for (int b = tid; b < N2; b += nth) sig_t[b] = sg[b];
  if (a.record)
    for (int q = tid; q < a.dwords; q += nth) a.record[(size_t)chain * a.dwords + q] = s_dec[q];
  if (tid == 0) {
    a.nacc[chain] = na;
    a.mu_done[chain] = mu_done;
    if (!a.gbond) {
      uint64_t x = (uint64_t)seed0 * (uint64_t)a.apow[N2];
      a.seed[chain] = (int32_t)(uint32_t)(x & 0x7fffffffull);
      a.acc_count[2 * chain] += (unsigned long long)tot;
      a.acc_count[2 * chain + 1] += (unsigned long long)N2;
    }
  }
}
