// Local link-flip sweep of one time slice with G held in registers (FP32 G,
// N = 256): the reference's LocalK_metro (src/localK.f90:45-173) in the order of
// LocalK_prop_L/R (:347-398), as k_flip, for the mixed mode at the size where
// the FP32 G of one chain (256 KiB) fits in the register file of one compute
// unit (512 KiB).
//
// One workgroup (8 wavefronts) per chain, thread k owns proposal k.  G is
// loaded once into the wavefronts' registers in the matrix cores' accumulator
// layout: wavefront w holds rows R0 = 64 (w / 2) .. R0+63 and columns
// C0 = 128 (w % 2) .. C0+127 as 4 x 8 tiles of 16 x 16, lane (li, lk) item q of
// tile (tr, tc) is G(R0 + 16 tr + 4 lk + q, C0 + 16 tc + li); a column of G is
// spread over four wavefronts and a row over two, so that copying one out
// takes few stores per lane.
// Per accepted proposal:
//   (E) every undecided proposal folds the last accepted update into its 2x2
//       block (registers) and is evaluated with its cached boson ratio; a
//       wavefront min and one LDS atomic give the first accepted one (the
//       serial decision, as in k_flip);
//   (X) the lanes holding columns p1, p2 and rows p1, p2 of G copy them to LDS;
//   (U) 2N threads build the update vectors u = G[:,P] Prod^{-1} and
//       v = Delta (I - G)[P,:] in FP64, minus the updates not yet in the
//       registers (at most two);
// with a barrier after each phase.
// Accepted updates are applied to the registers in pairs, one rank-4 product
// per tile on the FP32 matrix cores, issued after the extraction that follows
// their second member; a lone last update is applied at the end.  Nothing is
// left pending for the kinetic kernels: G is written back whole, with the
// exp(dtau mu) factors applied on load (left pass) or on store (right pass) as
// k_flip applies them.
#pragma once
#include "flip_kernels.hpp"

typedef float rg_f4 __attribute__((ext_vector_type(4)));

// LDS bytes of k_flip_rg for N sites
__host__ __device__ inline size_t flip_rg_lds_bytes(int N) {
  const size_t N2 = 2 * (size_t)N;
  size_t b = sizeof(double) * (4 * 4 * (size_t)N + 6 * N2 + 16);  // 4 update slots, Prod^-1 and det per bond, boson
  b += sizeof(float) * (2 + 8) * (size_t)N;  // columns p1, p2; rows p1, p2 as float4 per column
  b += sizeof(float) * 4 * 4 * (size_t)N;    // the update slots as the products' FP32 operands
  b += N2;                                   // current field
  return (b + 15) & ~(size_t)15;
}

// the rank-4 (both = true) or rank-2 update of the register tiles:
// G(r, c) -= sum_k u_k(r) v_k(c) over the vectors of slots sa (and sb), from the
// slots' FP32 copies (u_c(i) at 2i + c, -v_c(i) at 2N + 2i + c)
template <int TRW, int TCW>
__device__ __forceinline__ void rg_update(rg_f4 (&g)[TRW][TCW], const float *UF, int N, int R0, int C0, int li,
                                          int lk, int sa, int sb, bool both) {
  const int cmp = lk & 1;
  const float *U = UF + (size_t)4 * N * (lk < 2 ? sa : sb);
  const bool on = lk < 2 || both;
  float av[TRW], bv[TCW];
#pragma unroll
  for (int tr = 0; tr < TRW; ++tr) {
    const float t = U[2 * (R0 + 16 * tr + li) + cmp];
    av[tr] = on ? t : 0.f;  // u_k(row)
  }
#pragma unroll
  for (int tc = 0; tc < TCW; ++tc) {
    const float v = U[2 * N + 2 * (C0 + 16 * tc + li) + cmp];
    bv[tc] = on ? v : 0.f;  // -v_k(col)
  }
#pragma unroll
  for (int tr = 0; tr < TRW; ++tr)
#pragma unroll
    for (int tc = 0; tc < TCW; ++tc) g[tr][tc] = mfma_16x16x4<float>(av[tr], bv[tc], g[tr][tc]);
}

// boson ratio of flipping the field s_old of a bond with partner bonds bb and
// neighbouring-slice sum ss (src/fields.f90:66-117), from the current field sg
__device__ __forceinline__ double rg_boson(const int8_t *sg, const double *bos, const int (&bb)[6], int s_old, int ss) {
  const int s_new = -s_old;
  const int tn1 = s_new * (int)sg[bb[0]] * (int)sg[bb[1]] * (int)sg[bb[2]];
  const int tn2 = s_new * (int)sg[bb[3]] * (int)sg[bb[4]] * (int)sg[bb[5]];
  const int dS = 2 * (tn1 + tn2);
  const int dT = 2 * s_new * ss;
  return bos[(dS / 4 + 1) * 3 + (dT / 4 + 1)];
}

// sum over c of T_c(i) V_c(j) for the update vectors of one slot
__device__ __forceinline__ double rg_dot2(const double *T, int i, const double *V, int j) {
  return T[2 * i] * V[2 * j] + T[2 * i + 1] * V[2 * j + 1];
}

template <int NC>
__global__ void __launch_bounds__(512) k_flip_rg(FlipArgs<float> a) {
  static_assert(NC == 256, "register-resident G: 8 wavefronts x 64 x 128 blocks");
  constexpr int N = NC, N2 = 2 * NC, TRW = 4, TCW = 8, nth = 512;  // tiles per wavefront: 4 rows x 8 columns
  const int chain = blockIdx.x;
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63, li = lane & 15, lk = lane >> 4;
  const int R0 = 64 * (wave >> 1), C0 = 128 * (wave & 1);  // this wavefront's rows and columns of G
  const ChainConst *cp = a.cc + chain;
  const double ep = cp->ep, em = cp->em;
  float *G = a.G + (size_t)chain * N * N;
  int8_t *sigc = a.sig + (size_t)chain * a.T * N2;
  int8_t *sig_t = sigc + (size_t)(a.nt - 1) * N2;
  const int tp = (a.nt - 2 + a.T) % a.T, tn = a.nt % a.T;  // 0-based neighbour slices
  const int8_t *sig_p = sigc + (size_t)tp * N2, *sig_n = sigc + (size_t)tn * N2;

  extern __shared__ unsigned char smem_raw[];
  double *US = reinterpret_cast<double *>(smem_raw);   // 4 x 4N: update slots, u_c(i) at 2i + c,
                                                       //         v_c(i) at 2N + 2i + c
  double *o4 = US + (size_t)16 * N;                    // 2N x 6: Prod^{-1} (4) and det of an accepted candidate
  double *bos = o4 + (size_t)6 * N2;                   // 9 (16): boson ratio table
  float *col = reinterpret_cast<float *>(bos + 16);    // 2 x N : G(:,p1), G(:,p2)
  float *row4 = col + 2 * N;                           // 2 x N x 4: G(p,:) as the float4 of rows
                                                       //   4(p/4) .. 4(p/4)+3 per column
  float *USF = row4 + 8 * N;                           // 4 x 4N: the slots in FP32, v negated
  int8_t *sg = reinterpret_cast<int8_t *>(USF + 16 * N);  // 2N : current slice field
  __shared__ int s_kmin[2];

  const double Dp0 = cp->delta[0][0], Dp1 = cp->delta[0][1], Dp2 = cp->delta[0][2], Dp3 = cp->delta[0][3];
  const double Dm0 = cp->delta[1][0], Dm1 = cp->delta[1][1], Dm2 = cp->delta[1][2], Dm3 = cp->delta[1][3];
  const int mu = a.mu_on_read;

  const uint32_t seed0 = (uint32_t)a.seed[chain] & 0x7fffffffu;
  // this thread's proposal k = tid (2N = 512 threads): bond, sites, partner bonds,
  // field, random number (draw k+1 of the slice), 2x2 block, straight from global
  // memory so that these loads overlap the tables' copy into LDS
  const int k = tid;
  const int rb_ = a.descending ? (N2 - 1 - k) : k;
  const int rp1 = a.lat.bond_p[2 * rb_], rp2 = a.lat.bond_p[2 * rb_ + 1];
  int rbb[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) rbb[q] = a.lat.bond_bonds[6 * rb_ + q];
  const int rso = sig_t[rb_], rss = (int)sig_p[rb_] + (int)sig_n[rb_];
  const double rrn = (double)(uint32_t)(((uint64_t)seed0 * (uint64_t)a.apow[k + 1]) & 0x7fffffffull) / 2147483648.0;
  double r11 = (double)flip_read<float>(G, rp1, rp1, N, mu, ep, em);
  double r22 = (double)flip_read<float>(G, rp2, rp2, N, mu, ep, em);
  double r12 = (double)flip_read<float>(G, rp1, rp2, N, mu, ep, em);
  double r21 = (double)flip_read<float>(G, rp2, rp1, N, mu, ep, em);
  // G into the registers (with the mu factors of a left pass, as flip_read)
  const float fr1 = mu ? (float)ep : 1.f, fr2 = mu ? (float)em : 1.f;
  rg_f4 g[TRW][TCW];
#pragma unroll
  for (int tr = 0; tr < TRW; ++tr)
#pragma unroll
    for (int tc = 0; tc < TCW; ++tc) {
      rg_f4 x = *reinterpret_cast<const rg_f4 *>(G + R0 + 16 * tr + 4 * lk + (size_t)(C0 + 16 * tc + li) * N);
#pragma unroll
      for (int q = 0; q < 4; ++q) {  // times 1 (exact) without the factors
        x[q] = x[q] * fr1;
        x[q] = x[q] * fr2;
      }
      g[tr][tc] = x;
    }
  // lattice tables and the field of the slice into LDS
  for (int b = tid; b < N2; b += nth) sg[b] = sig_t[b];
  if (tid < 9) bos[tid] = cp->boson[tid];
  // the decision word: (2 k + (old field > 0)) << 16 | p1 << 8 | p2 for proposal k on the
  // bond of sites p1, p2 (N = 256: sites fit in 8 bits), so that its minimum over the
  // accepted proposals is the first one and carries the sites with it
  constexpr int KEY = 16, NONE = (2 * N2 + 2) << KEY;
  if (tid == 0) {
    s_kmin[0] = NONE;
    s_kmin[1] = NONE;
  }
  const uint32_t *frc = a.force ? a.force + (size_t)chain * a.dwords : nullptr;
  const int gb = a.gbond ? (a.gmask[(size_t)chain * a.T + (a.nt - 1)] ? a.gbond[chain] : -1) : -1;
  __syncthreads();
  double rbo = rg_boson(sg, bos, rbb, rso, rss);
  const double D0 = rso > 0 ? Dp0 : Dm0, D1 = rso > 0 ? Dp1 : Dm1;
  const double D2 = rso > 0 ? Dp2 : Dm2, D3 = rso > 0 ? Dp3 : Dm3;

  int start = 0, round = 0, tot = 0, last = -1, lastb = -1;
  bool mine = false;  // this thread's proposal was accepted
  int pq0 = -1, pq1 = -1;  // slots of the accepted updates not yet in the registers, in order
  while (true) {
    const int slot = round & 1;
    // (E) fold the last accepted update into this thread's undecided proposal, evaluate it
    int cand = NONE;
    if (k >= start) {
      if (last >= 0) {
        const double *TA = US + (size_t)4 * N * last, *VA = TA + 2 * N;
        const double d12 = rg_dot2(TA, rp1, VA, rp2), d21 = rg_dot2(TA, rp2, VA, rp1);
        const double e1 = rg_dot2(TA, rp1, VA, rp1), e2 = rg_dot2(TA, rp2, VA, rp2);
        r12 = r12 - d12;
        r21 = r21 - d21;
        r11 = r11 - e1;
        r22 = r22 - e2;
        // the boson ratio changes only when a partner bond flipped
        if (rbb[0] == lastb || rbb[1] == lastb || rbb[2] == lastb || rbb[3] == lastb || rbb[4] == lastb ||
            rbb[5] == lastb)
          rbo = rg_boson(sg, bos, rbb, rso, rss);
      }
      const double A11 = 1.0 - r11, A21 = 0.0 - r21;
      const double A12 = 0.0 - r12, A22 = 1.0 - r22;
      const double P11 = 1.0 + (D0 * A11 + D2 * A21);
      const double P21 = 0.0 + (D1 * A11 + D3 * A21);
      const double P12 = 0.0 + (D0 * A12 + D2 * A22);
      const double P22 = 1.0 + (D1 * A12 + D3 * A22);
      const double det = P11 * P22 - P12 * P21;
      const double ratio = (det * det) * rbo;
      const bool acc = a.gbond ? rb_ == gb
                               : (frc ? ((frc[k >> 5] >> (k & 31)) & 1u) != 0u : fabs(ratio) > rrn);
      if (acc) {
        // Prod^{-1} for the update, should this proposal be the first accepted one
        cand = ((2 * k + (rso > 0 ? 1 : 0)) << KEY) | (rp1 << 8) | rp2;
        const double rdet = 1.0 / det;
        double *o = o4 + 6 * rb_;
        o[0] = P22 * rdet;
        o[1] = (-P21) * rdet;
        o[2] = (-P12) * rdet;
        o[3] = P11 * rdet;
        o[4] = det;
      }
    }
    cand = wave_min_int(cand);
    if (lane == 0 && cand < NONE) atomicMin(&s_kmin[slot], cand);
    if (tid == 0) s_kmin[slot ^ 1] = NONE;
    __syncthreads();
    const int packed = __builtin_amdgcn_readfirstlane(s_kmin[slot]);
    if (packed >= NONE) break;
    const int kmin = packed >> (KEY + 1), s_old = ((packed >> KEY) & 1) ? 1 : -1;
    const int b = a.descending ? (N2 - 1 - kmin) : kmin;
    const int p1 = (packed >> 8) & 255, p2 = packed & 255;
    mine = mine || k == kmin;
    // (X) columns p1, p2 and rows p1, p2 of the register G into LDS
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int P = j ? p2 : p1;
      // column P: tile column (P >> 4) & 7 of the wavefronts with C0 = 128 (P >> 7), lanes li = P & 15
      if ((wave & 1) == (P >> 7) && li == (P & 15)) {
        float *dst = col + j * N + R0 + 4 * lk;
        switch ((P >> 4) & 7) {
#define RG_COL_CASE(t)                                                                      \
  case t:                                                                                   \
    _Pragma("unroll") for (int tr = 0; tr < TRW; ++tr) *reinterpret_cast<rg_f4 *>(dst + 16 * tr) = g[tr][t]; \
    break;
          RG_COL_CASE(0) RG_COL_CASE(1) RG_COL_CASE(2) RG_COL_CASE(3)
          RG_COL_CASE(4) RG_COL_CASE(5) RG_COL_CASE(6) RG_COL_CASE(7)
#undef RG_COL_CASE
          default: break;
        }
      }
      // row P: item P & 3 of tile row (P >> 4) & 3 of the wavefronts with R0 = 64 (P >> 6), lanes
      // lk = (P >> 2) & 3; the whole float4 is stored and the item picked by address when read
      if ((wave >> 1) == (P >> 6) && lk == ((P >> 2) & 3)) {
        float *dst = row4 + 4 * (j * N + C0 + li);
        switch ((P >> 4) & 3) {
#define RG_ROW_CASE(t)                                                                          \
  case t:                                                                                       \
    _Pragma("unroll") for (int tc = 0; tc < TCW; ++tc) *reinterpret_cast<rg_f4 *>(dst + 64 * tc) = g[t][tc]; \
    break;
          RG_ROW_CASE(0) RG_ROW_CASE(1) RG_ROW_CASE(2) RG_ROW_CASE(3)
#undef RG_ROW_CASE
          default: break;
        }
      }
    }
    // Prod^{-1} of the accepted proposal as its thread computed it when evaluating it
    const double D11 = s_old > 0 ? Dp0 : Dm0, D21 = s_old > 0 ? Dp1 : Dm1;
    const double D12 = s_old > 0 ? Dp2 : Dm2, D22 = s_old > 0 ? Dp3 : Dm3;
    const double *ob = o4 + 6 * b;
    const double Pi11 = ob[0], Pi21 = ob[1], Pi12 = ob[2], Pi22 = ob[3];
    if (a.gbond && tid == 0) {
      const double det = ob[4];
      double r = fabs(det * det);
      a.glr[2 * chain] = a.glr[2 * chain] + log(r > 1.0e-300 ? r : 1.0e-300);
    }
    if (tid == 0) {
      sg[b] = (int8_t)(-s_old);
      tot = tot + 1;
    }
    __syncthreads();
    // (U) this accept's update vectors in slot ns, from the extraction minus the
    // pending updates pq0, pq1 (FP64)
    const int q1 = p1 & 3, q2 = p2 & 3;
    const double *T0 = pq0 >= 0 ? US + (size_t)4 * N * pq0 : nullptr, *V0 = pq0 >= 0 ? T0 + 2 * N : nullptr;
    const double *T1 = pq1 >= 0 ? US + (size_t)4 * N * pq1 : nullptr, *V1 = pq1 >= 0 ? T1 + 2 * N : nullptr;
    // G(i, p1), G(i, p2) and G(p1, i), G(p2, i) of the current G at site i
    auto colv = [&](int i, double &c1, double &c2) {
      c1 = (double)col[i];
      c2 = (double)col[N + i];
      if (pq0 >= 0) {
        c1 = c1 - rg_dot2(T0, i, V0, p1);
        c2 = c2 - rg_dot2(T0, i, V0, p2);
      }
      if (pq1 >= 0) {
        c1 = c1 - rg_dot2(T1, i, V1, p1);
        c2 = c2 - rg_dot2(T1, i, V1, p2);
      }
    };
    auto rowv = [&](int i, double &r1, double &r2) {
      r1 = (double)row4[4 * i + q1];
      r2 = (double)row4[4 * (N + i) + q2];
      if (pq0 >= 0) {
        r1 = r1 - rg_dot2(T0, p1, V0, i);
        r2 = r2 - rg_dot2(T0, p2, V0, i);
      }
      if (pq1 >= 0) {
        r1 = r1 - rg_dot2(T1, p1, V1, i);
        r2 = r2 - rg_dot2(T1, p2, V1, i);
      }
    };
    auto u_at = [&](int i, double &u0, double &u1) {
      double c1, c2;
      colv(i, c1, c2);
      u0 = c1 * Pi11 + c2 * Pi21;
      u1 = c1 * Pi12 + c2 * Pi22;
    };
    auto v_at = [&](int i, double &v0, double &v1) {
      double r1, r2;
      rowv(i, r1, r2);
      v0 = -(D11 * r1) - D12 * r2;
      v1 = -(D21 * r1) - D22 * r2;
      if (i == p1) {
        v0 = v0 + D11;
        v1 = v1 + D21;
      }
      if (i == p2) {
        v0 = v0 + D12;
        v1 = v1 + D22;
      }
    };
    const int ns = round & 3;
    double *TN = US + (size_t)4 * N * ns;
    float *TF = USF + (size_t)4 * N * ns;
    // the two pending updates into the registers (the extraction above has read them):
    // half of the wavefronts issue their products before this phase's vector work, half
    // after it, so that each SIMD's matrix core starts on one wavefront's products while
    // its other wavefront computes (early: w and w ^ 1, and w and w ^ 4, differ)
    const bool early = ((wave ^ (wave >> 2)) & 1) == 0;
    if (pq1 >= 0 && early) rg_update<TRW, TCW>(g, USF, N, R0, C0, li, lk, pq0, pq1, true);
    if (tid < N) {
      double u0, u1;
      u_at(tid, u0, u1);
      TN[2 * tid] = u0;
      TN[2 * tid + 1] = u1;
      TF[2 * tid] = (float)u0;
      TF[2 * tid + 1] = (float)u1;
    } else {
      double v0, v1;
      v_at(tid - N, v0, v1);
      TN[2 * N + 2 * (tid - N)] = v0;
      TN[2 * N + 2 * (tid - N) + 1] = v1;
      TF[2 * N + 2 * (tid - N)] = (float)(-v0);
      TF[2 * N + 2 * (tid - N) + 1] = (float)(-v1);
    }
    if (pq1 >= 0 && !early) rg_update<TRW, TCW>(g, USF, N, R0, C0, li, lk, pq0, pq1, true);
    __syncthreads();
    if (pq1 >= 0) {
      pq0 = ns;
      pq1 = -1;
    } else if (pq0 >= 0) {
      pq1 = ns;
    } else {
      pq0 = ns;
    }
    last = ns;
    lastb = b;
    start = kmin + 1;
    round = round + 1;
  }
  if (pq0 >= 0) rg_update<TRW, TCW>(g, USF, N, R0, C0, li, lk, pq0, pq1 >= 0 ? pq1 : pq0, pq1 >= 0);
  // G back to memory (right pass: with the mu factors, as k_flip's final flush)
  const float fw1 = a.final_flush == 1 ? (float)ep : 1.f, fw2 = a.final_flush == 1 ? (float)em : 1.f;
#pragma unroll
  for (int tr = 0; tr < TRW; ++tr)
#pragma unroll
    for (int tc = 0; tc < TCW; ++tc) {
      rg_f4 x = g[tr][tc];
#pragma unroll
      for (int q = 0; q < 4; ++q) {  // times 1 (exact) without the factors
        x[q] = x[q] * fw1;
        x[q] = x[q] * fw2;
      }
      *reinterpret_cast<rg_f4 *>(G + R0 + 16 * tr + 4 * lk + (size_t)(C0 + 16 * tc + li) * N) = x;
    }
  for (int bb = tid; bb < N2; bb += nth) sig_t[bb] = sg[bb];
  if (a.record) {  // proposal k = tid: bit k of the decision words, one 64-bit ballot per wavefront
    const unsigned long long bits = __ballot(mine);
    if (lane == 0) {
      a.record[(size_t)chain * a.dwords + 2 * wave] = (uint32_t)bits;
      a.record[(size_t)chain * a.dwords + 2 * wave + 1] = (uint32_t)(bits >> 32);
    }
  }
  if (tid == 0) {
    a.nacc[chain] = 0;
    a.mu_done[chain] = mu;
    if (!a.gbond) {
      uint64_t x = (uint64_t)seed0 * (uint64_t)a.apow[N2];
      a.seed[chain] = (int32_t)(uint32_t)(x & 0x7fffffffull);
      a.acc_count[2 * chain] += (unsigned long long)tot;
      a.acc_count[2 * chain + 1] += (unsigned long long)N2;
    }
  }
}
