// dev_filter.h -- the per-candidate-pair body of the banded u8 filter (used by k_filter and
// k_fused).  Semantics and the two exact rewrites are described in k_filter.hip.
#pragma once
#include "storm_dev.h"

namespace storm {

typedef short s2 __attribute__((ext_vector_type(2)));
typedef unsigned short u2 __attribute__((ext_vector_type(2)));

// One row of a candidate pair: P[d] holds, for diagonal d, the mismatch bits of 16 consecutive
// rows of candidate a (low half) and candidate b (high half); this row is bit t of each half.
// FAST (gap open - gap extend == match, gap open >= gap extend: the default scores): with
// X = corner + m - (m + x) * mismatch - (open - ext) = corner - (m + x) * mismatch, the gap state
// max(subs(I, ext), subs(M, open)) is subs(max(I, X), ext) and the cell's score X + (open - ext),
// one instruction fewer per cell pair (M = max(X + open - ext, I) >= I, so subs(M, open) adds
// nothing to max(I - ext, X + open - ext - open)).
template <int D, bool MASKED, bool FAST>
__device__ __forceinline__ void filter_row(s2 (&M)[D], u2 (&G)[D], u2& best, const uint32_t (&P)[D], int t,
                                           s2 NEG, s2 MAT, u2 EXT, u2 OPN, u2 vm) {
#pragma unroll
  for (int d = 0; d < D; ++d) {
    uint32_t nzb = (P[d] >> t) & 0x00010001u;
    s2 nz = __builtin_bit_cast(s2, nzb);
    u2 I;
    if (d == 0) I = G[1];
    else if (d == D - 1) I = G[D - 2];
    else I = __builtin_elementwise_max(G[d + 1], G[d - 1]);   // top (old row), left (new row)
    const s2 Is = __builtin_bit_cast(s2, I);
    s2 Mv;
    if (FAST) {
      const s2 X = M[d] + nz * NEG;                         // corner + (match ? m : -x) - (open - ext)
      Mv = __builtin_elementwise_max(X + MAT, Is);
      G[d] = __builtin_elementwise_sub_sat(__builtin_bit_cast(u2, __builtin_elementwise_max(Is, X)), EXT);
    } else {
      const s2 Dv = M[d] + nz * NEG + MAT;                  // corner + (match ? m : -x)
      Mv = __builtin_elementwise_max(Dv, Is);
      G[d] = __builtin_elementwise_max(__builtin_elementwise_sub_sat(I, EXT),
                                       __builtin_elementwise_sub_sat(__builtin_bit_cast(u2, Mv), OPN));
    }
    const u2 Mu = __builtin_bit_cast(u2, Mv);
    M[d] = Mv;
    if (MASKED) best = __builtin_elementwise_max(best, Mu & vm);
    else best = __builtin_elementwise_max(best, Mu);
  }
}


struct FilterConst {
  s2 NEG, MAT;
  u2 EXT, OPN;
  int thr;
  bool fast;   // open - ext == m and open >= ext: filter_row's FAST form
};

__device__ __forceinline__ FilterConst filter_const(int mt, int mm, int op, int ex, int thr) {
  FilterConst k;
  k.NEG = (s2){(short)-(mt + mm), (short)-(mt + mm)};
  k.MAT = (s2){(short)mt, (short)mt};
  k.EXT = (u2){(unsigned short)ex, (unsigned short)ex};
  k.OPN = (u2){(unsigned short)op, (unsigned short)op};
  k.thr = thr;
  k.fast = (op - ex == mt) && op >= ex;
  return k;
}

// Two candidates: read planes pa/pb of lengths La/Lb, windows starting at genome codes xa/xb
// (already offset by GPAD).  Returns the reference's u8 running maxima in sa/sb.
template <int W, bool FAST>
__device__ __forceinline__ void filter_pair_f(const GenomeDev& g, const uint32_t* pa, const uint32_t* pb, int La, int Lb,
                                              int64_t xa, int64_t xb, const FilterConst& k, int& sa, int& sb) {
  constexpr int D = 2 * W;
  const int Lmin = La < Lb ? La : Lb, Lmax = La < Lb ? Lb : La;
  const u2 vm = {(unsigned short)(La > Lb ? 0xFFFF : 0), (unsigned short)(Lb > La ? 0xFFFF : 0)};
  s2 M[D];
  u2 G[D];
#pragma unroll
  for (int d = 0; d < D; ++d) { M[d] = (s2){0, 0}; G[d] = (u2){0, 0}; }
  u2 best = {0, 0};
  for (int blk = 0; blk * 32 < Lmax; ++blk) {
    // genome planes: 64 bits starting at window code 32*blk of each candidate
    const int64_t ba = xa + 32 * blk, bb = xb + 32 * blk;
    uint32_t A0l, A0h, A1l, A1h, B0l, B0h, B1l, B1h;
    gbit_pair(g.lh, ba, A0l, A0h);
    gbit_pair(g.lh, ba + 32, A1l, A1h);
    gbit_pair(g.lh, bb, B0l, B0h);
    gbit_pair(g.lh, bb + 32, B1l, B1h);
    const uint32_t ral = pa[blk], rah = pa[RW + blk], rbl = pb[blk], rbh = pb[RW + blk];
    // mismatch bits of the block's 32 rows per diagonal: bit t of diagonal d compares read code
    // 32*blk + t with genome code ba + t + d (one funnel shift of each plane per diagonal, not
    // per row); Q holds rows 0-15 and R rows 16-31, candidate a in the low half, b in the high
    uint32_t Q[D], R[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const uint32_t ma = (__builtin_amdgcn_alignbit(A1l, A0l, d) ^ ral) | (__builtin_amdgcn_alignbit(A1h, A0h, d) ^ rah);
      const uint32_t mb = (__builtin_amdgcn_alignbit(B1l, B0l, d) ^ rbl) | (__builtin_amdgcn_alignbit(B1h, B0h, d) ^ rbh);
      Q[d] = __builtin_amdgcn_perm(mb, ma, 0x05040100u);
      R[d] = __builtin_amdgcn_perm(mb, ma, 0x07060302u);
    }
    // each half: its rows below both lengths, then the rows only the longer candidate has (its
    // partner's maximum frozen), as two loops, so no row branches between the two forms
    const int r0 = 32 * blk;
    const int rows = (Lmax - r0) < 32 ? (Lmax - r0) : 32;
    const int n0 = rows < 16 ? rows : 16, n1 = rows - n0;
    const int u0 = min(max(Lmin - r0, 0), n0), u1 = min(max(Lmin - r0 - 16, 0), n1);
    for (int t = 0; t < u0; ++t) filter_row<D, false, FAST>(M, G, best, Q, t, k.NEG, k.MAT, k.EXT, k.OPN, vm);
    for (int t = u0; t < n0; ++t) filter_row<D, true, FAST>(M, G, best, Q, t, k.NEG, k.MAT, k.EXT, k.OPN, vm);
    for (int t = 0; t < u1; ++t) filter_row<D, false, FAST>(M, G, best, R, t, k.NEG, k.MAT, k.EXT, k.OPN, vm);
    for (int t = u1; t < n1; ++t) filter_row<D, true, FAST>(M, G, best, R, t, k.NEG, k.MAT, k.EXT, k.OPN, vm);
  }
  sa = best.x > 255 ? 255 : best.x;
  sb = best.y > 255 ? 255 : best.y;
}

template <int W>
__device__ __forceinline__ void filter_pair(const GenomeDev& g, const uint32_t* pa, const uint32_t* pb, int La, int Lb,
                                            int64_t xa, int64_t xb, const FilterConst& k, int& sa, int& sb) {
  if (k.fast) filter_pair_f<W, true>(g, pa, pb, La, Lb, xa, xb, k, sa, sb);
  else filter_pair_f<W, false>(g, pa, pb, La, Lb, xa, xb, k, sa, sb);
}

}  // namespace storm
