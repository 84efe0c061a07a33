// k_filter.hip -- the banded u8 filter (stage c) for gfx950.
//
// Semantics: one candidate of alignment_sse2__align_{mono,pair,quad,octa}
// (alignment_simd.c:3731-4511, geometry run.c:109-146, threshold :3643-3645): a local
// Smith-Waterman over rows i in [0, L) and diagonals delta in [0, 2W) of the genome window that
// starts at p (genome code j = p + i + delta), unsigned 8-bit saturating scores, one merged gap
// state, zero boundaries; the candidate passes when its maximum exceeds min(z, 254).
//
// gfx950 mapping.  Each lane owns TWO candidates, one per 16-bit half of every VGPR, and walks
// their rows; the 2W diagonals of a row are unrolled in registers, so no cross-lane traffic is
// needed.  Per pair of cells the chain is ~10 packed VALU ops (v_pk_mad_u16, v_pk_add_u16,
// v_pk_max_u16/i16, v_pk_sub_u16 clamp).  Two rewrites keep it short and exact:
//   * the gap state is carried as G = max(subs(I, ext), subs(M, open)) per cell, so a cell's
//     I is max(G_top, G_left) -- one op instead of four;
//   * the upper u8 clamp is dropped: scores are kept unclamped in 16 bits and the final
//     maximum is clamped to 255.  The first cell that would saturate makes the clamped maximum
//     255 and every cell before it is unchanged, so min(255, max) is the reference's result.
// Read and genome codes arrive as bit planes; a row's 2W match/mismatch bits for one candidate
// cost 7 ops (two funnel shifts of the genome planes, two sign-extended read bits, xor/xor/or).
#include "storm_kernels.h"

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

template <int W>
__global__ __launch_bounds__(256) void k_filter(FilterArgs a) {
  constexpr int D = 2 * W;
  const int64_t pair = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t c0 = 2 * pair;
  if (c0 >= a.n) return;
  const int64_t c1 = (c0 + 1 < a.n) ? c0 + 1 : c0;
  const int ra = a.cand_read[c0], rb = a.cand_read[c1];
  const int La = a.rlen[ra >> a.len_shift], Lb = a.rlen[rb >> a.len_shift];
  const uint32_t* pa = a.rplanes + (int64_t)ra * (2 * RW);
  const uint32_t* pb = a.rplanes + (int64_t)rb * (2 * RW);
  const int64_t xa = (int64_t)a.cand_pos[c0] + GPAD, xb = (int64_t)a.cand_pos[c1] + GPAD;
  const int Lmin = La < Lb ? La : Lb, Lmax = La < Lb ? Lb : La;

  const s2 NEG = {(short)-(a.mt + a.mm), (short)-(a.mt + a.mm)};
  const s2 MAT = {(short)a.mt, (short)a.mt};
  const u2 EXT = {(unsigned short)a.ex, (unsigned short)a.ex};
  const u2 OPN = {(unsigned short)a.op, (unsigned short)a.op};
  const u2 vm = {(unsigned short)(La > Lb ? 0xFFFF : 0), (unsigned short)(Lb > La ? 0xFFFF : 0)};

  s2 M[D];
  u2 G[D];
#pragma unroll
  for (int d = 0; d < D; ++d) { M[d] = (s2){0, 0}; G[d] = (u2){0, 0}; }
  u2 best = {0, 0};

  for (int blk = 0; blk * 32 < Lmax; ++blk) {
    // genome planes: 64 bits starting at window code 32*blk of each candidate
    const int64_t ba = xa + 32 * blk, bb = xb + 32 * blk;
    const uint32_t A0l = gbit_word(a.g.lo, ba), A1l = gbit_word(a.g.lo, ba + 32);
    const uint32_t A0h = gbit_word(a.g.hi, ba), A1h = gbit_word(a.g.hi, ba + 32);
    const uint32_t B0l = gbit_word(a.g.lo, bb), B1l = gbit_word(a.g.lo, bb + 32);
    const uint32_t B0h = gbit_word(a.g.hi, bb), B1h = gbit_word(a.g.hi, bb + 32);
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
      if (32 * blk + t < Lmin) filter_row<D, false>(M, G, best, mlo, mhi, NEG, MAT, EXT, OPN, vm);
      else filter_row<D, true>(M, G, best, mlo, mhi, NEG, MAT, EXT, OPN, vm);
    }
  }
  const int sa = best.x > 255 ? 255 : best.x;
  const int sb = best.y > 255 ? 255 : best.y;
  if (a.score) {
    a.score[c0] = (uint8_t)sa;
    if (c1 != c0) a.score[c1] = (uint8_t)sb;
  }
  if (a.pass) {
    a.pass[c0] = (uint8_t)(sa > a.thr);
    if (c1 != c0) a.pass[c1] = (uint8_t)(sb > a.thr);
  }
}

// host-side launcher, W in {2, 4, 8, 16}
hipError_t launch_filter(const FilterArgs& a, int W, hipStream_t st) {
  const int64_t pairs = (a.n + 1) / 2;
  if (pairs == 0) return hipSuccess;
  const int bs = 256;
  const dim3 grid((unsigned)((pairs + bs - 1) / bs));
  switch (W) {
    case 2: hipLaunchKernelGGL(k_filter<2>, grid, dim3(bs), 0, st, a); break;
    case 4: hipLaunchKernelGGL(k_filter<4>, grid, dim3(bs), 0, st, a); break;
    case 8: hipLaunchKernelGGL(k_filter<8>, grid, dim3(bs), 0, st, a); break;
    case 16: hipLaunchKernelGGL(k_filter<16>, grid, dim3(bs), 0, st, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace storm
