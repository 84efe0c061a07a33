// dev_filter.h -- the per-candidate-pair body of the banded u8 filter (used by k_filter and
// k_fused).  Semantics and the two exact rewrites are described in k_filter.hip.
#pragma once
#include "storm_dev.h"

namespace storm {

typedef short s2 __attribute__((ext_vector_type(2)));
typedef unsigned short u2 __attribute__((ext_vector_type(2)));

template <int D, bool MASKED>
__device__ __forceinline__ void filter_row(s2 (&M)[D], u2 (&G)[D], u2& best, uint32_t mlo, uint32_t mhi,
                                           s2 NEG, s2 MAT, u2 EXT, u2 OPN, u2 vm) {
#pragma unroll
  for (int d = 0; d < D; ++d) {
    uint32_t nzb = ((d < 16 ? mlo : mhi) >> (d & 15)) & 0x00010001u;
    s2 nz = __builtin_bit_cast(s2, nzb);
    s2 Dv = M[d] + nz * NEG + MAT;                         // corner + (match ? m : -x)
    u2 I;
    if (d == 0) I = G[1];
    else if (d == D - 1) I = G[D - 2];
    else I = __builtin_elementwise_max(G[d + 1], G[d - 1]);   // top (old row), left (new row)
    s2 Mv = __builtin_elementwise_max(Dv, __builtin_bit_cast(s2, I));
    u2 Mu = __builtin_bit_cast(u2, Mv);
    G[d] = __builtin_elementwise_max(__builtin_elementwise_sub_sat(I, EXT), __builtin_elementwise_sub_sat(Mu, OPN));
    M[d] = Mv;
    if (MASKED) best = __builtin_elementwise_max(best, Mu & vm);
    else best = __builtin_elementwise_max(best, Mu);
  }
}


struct FilterConst {
  s2 NEG, MAT;
  u2 EXT, OPN;
  int thr;
};

__device__ __forceinline__ FilterConst filter_const(int mt, int mm, int op, int ex, int thr) {
  FilterConst k;
  k.NEG = (s2){(short)-(mt + mm), (short)-(mt + mm)};
  k.MAT = (s2){(short)mt, (short)mt};
  k.EXT = (u2){(unsigned short)ex, (unsigned short)ex};
  k.OPN = (u2){(unsigned short)op, (unsigned short)op};
  k.thr = thr;
  return k;
}

// Two candidates: read planes pa/pb of lengths La/Lb, windows starting at genome codes xa/xb
// (already offset by GPAD).  Returns the reference's u8 running maxima in sa/sb.
template <int W>
__device__ __forceinline__ void filter_pair(const GenomeDev& g, const uint32_t* pa, const uint32_t* pb, int La, int Lb,
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
    const int rows = (Lmax - 32 * blk) < 32 ? (Lmax - 32 * blk) : 32;
    for (int t = 0; t < rows; ++t) {
      const uint32_t gal = __builtin_amdgcn_alignbit(A1l, A0l, t), gah = __builtin_amdgcn_alignbit(A1h, A0h, t);
      const uint32_t gbl = __builtin_amdgcn_alignbit(B1l, B0l, t), gbh = __builtin_amdgcn_alignbit(B1h, B0h, t);
      const uint32_t sal = (uint32_t)__builtin_amdgcn_sbfe((int)ral, t, 1), sah = (uint32_t)__builtin_amdgcn_sbfe((int)rah, t, 1);
      const uint32_t sbl = (uint32_t)__builtin_amdgcn_sbfe((int)rbl, t, 1), sbh = (uint32_t)__builtin_amdgcn_sbfe((int)rbh, t, 1);
      const uint32_t ma = (gal ^ sal) | (gah ^ sah);      // bit delta = mismatch on diagonal delta
      const uint32_t mb = (gbl ^ sbl) | (gbh ^ sbh);
      const uint32_t mlo = __builtin_amdgcn_perm(mb, ma, 0x05040100u);   // [ma.lo16 | mb.lo16]
      const uint32_t mhi = (D > 16) ? __builtin_amdgcn_perm(mb, ma, 0x07060302u) : 0u;
      if (32 * blk + t < Lmin) filter_row<D, false>(M, G, best, mlo, mhi, k.NEG, k.MAT, k.EXT, k.OPN, vm);
      else filter_row<D, true>(M, G, best, mlo, mhi, k.NEG, k.MAT, k.EXT, k.OPN, vm);
    }
  }
  sa = best.x > 255 ? 255 : best.x;
  sb = best.y > 255 ? 255 : best.y;
}

}  // namespace storm
