// dev_align2.h -- the score pass of the banded DP for two records per lane, one in each 16-bit
// half of every register (band bounds IM <= 7: a row's 2IM+1 <= 15 columns of both records fit the
// halves of one 32-bit bit plane).
//
// Same recurrence as DPLane::row (alignment.c:465-604, color state alignment.c:533-599), written
// with packed 16-bit arithmetic (v_pk_add_i16 with saturation, v_pk_max_i16) and per-half masks
// (0xFFFF / 0) selected by bitwise ops, so one instruction stream advances both records.
//
// Saturation is exact for what the score pass decides (score > t and S >= t): scores of cells
// reached without a masked cell or the band border stay within [-2600, 2100] (at most 287 steps of
// at most -9 or +8), while any value through a masked cell or the border carries MIN_SCORE
// (-16383) and stays below -14000 in 32-bit and in saturated 16-bit arithmetic alike.  Comparisons
// between such a value and a normal one agree, so every normal cell, its `from`, indels and color
// state are the same as in 32 bits; a row or a last row without a normal cell gives S or score
// below -14000 + 1280 < t either way.
#pragma once
#include "dev_align.h"
#include "dev_filter.h"

namespace storm {

__device__ __forceinline__ uint32_t pu(s2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ s2 ps(uint32_t v) { return __builtin_bit_cast(s2, v); }
__device__ __forceinline__ u2 pw(uint32_t v) { return __builtin_bit_cast(u2, v); }
__device__ __forceinline__ uint32_t pwu(u2 v) { return __builtin_bit_cast(uint32_t, v); }
__host__ __device__ constexpr uint32_t pk(int v) { return ((uint32_t)(uint16_t)(int16_t)v) * 0x10001u; }

// Per-half masks (0xFFFF where true).  Written as packed VALU instructions: left to the compiler,
// these patterns become a compare and a select per half, which undoes the packing.
__device__ __forceinline__ uint32_t m_gt(uint32_t a, uint32_t b) {   // signed a > b: sign of sat(b - a)
  uint32_t d, m;
  asm("v_pk_sub_i16 %0, %1, %2 clamp" : "=v"(d) : "v"(b), "v"(a));
  // op_sel_hi:[0,1]: the high half also shifts by the constant's low half (15)
  asm("v_pk_ashrrev_i16 %0, 15, %1 op_sel_hi:[0,1]" : "=v"(m) : "v"(d));
  return m;
}
__device__ __forceinline__ uint32_t m_zero(uint32_t x) {   // x == 0: min(x, 1) - 1
  const uint32_t one = 0x00010001u;
  uint32_t t, m;
  asm("v_pk_min_u16 %0, %1, %2" : "=v"(t) : "v"(x), "v"(one));
  asm("v_pk_sub_u16 %0, %1, %2" : "=v"(m) : "v"(t), "v"(one));
  return m;
}
__device__ __forceinline__ uint32_t psel(uint32_t m, uint32_t x, uint32_t y) { return (x & m) | (y & ~m); }
__device__ __forceinline__ uint32_t padd(uint32_t a, uint32_t b) {
  return pu(__builtin_elementwise_add_sat(ps(a), ps(b)));
}
__device__ __forceinline__ uint32_t psub16(uint32_t a, uint32_t b) { return pwu(pw(a) - pw(b)); }   // per half
__device__ __forceinline__ uint32_t pmad(uint32_t a, uint32_t b, uint32_t c) { return pwu(pw(a) * pw(b) + pw(c)); }
__device__ __forceinline__ uint32_t pmax(uint32_t a, uint32_t b) { return pu(__builtin_elementwise_max(ps(a), ps(b))); }
__device__ __forceinline__ uint32_t pmin(uint32_t a, uint32_t b) { return pu(__builtin_elementwise_min(ps(a), ps(b))); }
// bit c of each half as a mask
template <int C>
__device__ __forceinline__ uint32_t pbit(uint32_t p) {
  const u2 x = pw(p) << (u2){(unsigned short)(15 - C), (unsigned short)(15 - C)};
  return pu(__builtin_bit_cast(s2, x) >> (s2){15, 15});
}
// bit C of each half moved to bit P of the half, other bits cleared
template <int C, int P>
__device__ __forceinline__ uint32_t pmove(uint32_t p) {
  u2 x = pw(p);
  if constexpr (C > P) x = x >> (u2){(unsigned short)(C - P), (unsigned short)(C - P)};
  else if constexpr (C < P) x = x << (u2){(unsigned short)(P - C), (unsigned short)(P - C)};
  return pwu(x) & (0x10001u << P);
}

// One genome window and read of the pair (one half).
struct DPHalf {
  uint32_t gl[3], gh[3], gm[3];
  int64_t wi;
  int sh, L, win;
  const uint32_t* pl;
  const uint32_t* qp;
  uint32_t rlo, rhi, rq;

  __device__ void begin(const GenomeDev& g, const ReadsDev& r, const DPConst& k, int rs, int h) {
    L = r.len[rs >> 1];
    win = (h - k.I > -k.d) ? h - k.I : -k.d;
    pl = read_planes(r, rs);
    qp = r.qual + (int64_t)rs * QW;
    const int64_t wbase = (int64_t)win + GPAD;
    wi = wbase >> 5;
    sh = (int)(wbase & 31);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const uint2 p = g.lh[wi + q];
      gl[q] = p.x; gh[q] = p.y; gm[q] = g.mask[wi + q];
    }
    rlo = pl[0];
    rhi = pl[RW];
    rq = r.has_qual ? qp[0] : 0xffffffffu;
  }
  // advance to row i (i >= 1): genome windows by one code, read and quality words as needed
  __device__ void advance(const GenomeDev& g, int i, int has_qual) {
    if (++sh == 32) {
      sh = 0;
      ++wi;
      gl[0] = gl[1]; gh[0] = gh[1]; gm[0] = gm[1];
      gl[1] = gl[2]; gh[1] = gh[2]; gm[1] = gm[2];
      const uint2 p = g.lh[wi + 2];
      gl[2] = p.x;
      gh[2] = p.y;
      gm[2] = g.mask[wi + 2];
    }
    if ((i & 31) == 0) { rlo = pl[i >> 5]; rhi = pl[RW + (i >> 5)]; }
    if ((i & 15) == 0 && has_qual) rq = qp[i >> 4];
  }
  __device__ uint32_t rcode(int i) const { return ((rlo >> (i & 31)) & 1u) | (((rhi >> (i & 31)) & 1u) << 1); }
  __device__ uint32_t qlevel(int i) const { return (rq >> ((i & 15) << 1)) & 3u; }
  __device__ uint32_t wlo() const { return __builtin_amdgcn_alignbit(gl[1], gl[0], sh); }
  __device__ uint32_t whi() const { return __builtin_amdgcn_alignbit(gh[1], gh[0], sh); }
  __device__ uint32_t wmk() const { return __builtin_amdgcn_alignbit(gm[1], gm[0], sh); }
};

// The traceback walk of one record (alignment.c:642-734) from (L-1, bc) over its direction words
// (nibble c & 7 of word (i * DW + (c >> 3)) at tbw[64 * 2 * (i * DW + (c >> 3))]); codes to tb.
template <int IM, bool COLOR>
__device__ __forceinline__ void dp_walk(const GenomeDev& g, const uint32_t* pl, int L, int win, int NC, int bc, const uint32_t* tbw,
                        uint8_t* tb, AlignRec& o) {
  constexpr int DW = (2 * IM + 1 + 7) / 8;
  int i = L - 1, c = bc, n = 0, color = 0;
  while (i >= 0 && c >= 0 && c < NC) {
    const int j = i + c;
    o.best_i0 = (int16_t)i;
    o.best_j0 = (int16_t)j;
    const int cell = (int)(tbw[128 * (i * DW + (c >> 3))] >> ((c & 7) * 4)) & 0xF;
    const int from = cell & 3;
    const int ri = rcode(pl, i);
    int code;
    if (from == 2) code = 0xC | ri;             // insertion in the read
    else if (from == 1) code = 0x1;             // deletion
    else {
      const int gj = gcode(g, (int64_t)win + j);
      if (ri != gj) {
        if (!COLOR) code = 0x4 | ri;
        else if (color || (i != 0 && (cell & 4))) { code = 0x8 | ri; color ^= (ri ^ gj); }
        else code = 0x4 | ri;
      } else code = 0x3;
    }
    // TRACEBACK_SEQ_SET (alignment.h:83): even index -> high nibble
    if (n & 1) tb[n >> 1] = (uint8_t)((tb[n >> 1] & 0xF0) | code);
    else tb[n >> 1] = (uint8_t)(code << 4);
    ++n;
    if (from == 3) { --i; }
    else if (from == 2) { --i; ++c; }
    else { --c; }
  }
  o.tb_len = (int16_t)n;
}

// Records (rsA, hA) and (rsB, hB), one per 16-bit half: the same score (MIN_SCORE when the pass
// stops because S < t) and survival bound S as dp_one<IM, COLOR, false>.  With TB, the rows'
// directions go to the lane's words (record A at tbw[128 k], B at tbw[128 k + 64], k = i * DW + w;
// the 64 lanes of a wavefront interleave theirs) and a record that can be accepted (score > t,
// S >= t) gets best_j and its traceback walk to tbA / tbB, as dp_one<IM, COLOR, true>.
template <int IM, bool COLOR, bool TB = false>
__device__ __forceinline__ void dp_score_pair(const GenomeDev& g, const ReadsDev& r, const ScoreDev& scd, int rsA, int hA, int rsB,
                              int hB, AlignRec& outA, AlignRec& outB, uint32_t* tbw = nullptr, uint8_t* tbA = nullptr,
                              uint8_t* tbB = nullptr) {
  static_assert(IM <= 7, "two records per lane need 2IM+1 <= 16 columns");
  constexpr int NCM = 2 * IM + 1;
  constexpr int DW = (NCM + 7) / 8;
  const DPConst k(scd);
  const int NC = k.NC;
  DPHalf A, B;
  A.begin(g, r, k, rsA, hA);
  B.begin(g, r, k, rsB, hB);
  const int has_qual = r.has_qual;
  // pair scores by quality level as 16-bit table entries: level q at bytes 2q, 2q+1
  const uint32_t tm01 = (uint32_t)(uint16_t)k.pm0 | ((uint32_t)(uint16_t)k.pm1 << 16);
  const uint32_t tm23 = (uint32_t)(uint16_t)k.pm2 | ((uint32_t)(uint16_t)k.pm3 << 16);
  const uint32_t tx01 = (uint32_t)(uint16_t)k.px0 | ((uint32_t)(uint16_t)k.px1 << 16);
  const uint32_t tx23 = (uint32_t)(uint16_t)k.px2 | ((uint32_t)(uint16_t)k.px3 << 16);
  const uint32_t MIN2 = pk(MIN_SCORE), OPEN2 = pk(k.gap_open), EXT2 = pk(k.gap_ext), I4_2 = pk(k.I4);
  const uint32_t BORDER2 = pk(m_make(3, k.I, 0, 0, 0));
  const uint32_t DGAP2 = psub16(OPEN2, EXT2);
  uint32_t sc[NCM], meta[NCM];
  // row 0 (alignment.c:326-360)
  {
    const uint32_t r0 = A.rcode(0) | (B.rcode(0) << 16), q0 = A.qlevel(0) | (B.qlevel(0) << 16);
    const uint32_t qsel = (q0 << 1) | (q0 << 9) | 0x01000100u;
    const uint32_t s_m = __builtin_amdgcn_perm(tm23, tm01, qsel), s_x = __builtin_amdgcn_perm(tx23, tx01, qsel);
    const uint32_t Plo = __builtin_amdgcn_perm(B.wlo(), A.wlo(), 0x05040100u);
    const uint32_t Phi = __builtin_amdgcn_perm(B.whi(), A.whi(), 0x05040100u);
    const uint32_t Pm = __builtin_amdgcn_perm(B.wmk(), A.wmk(), 0x05040100u);
    const uint32_t X0 = Plo ^ psub16(0, r0 & 0x00010001u), X1 = Phi ^ psub16(0, (r0 >> 1) & 0x00010001u);
    const uint32_t MIS = X0 | X1;
    const uint32_t q12 = q0 << 12;
#pragma unroll
    for (int c = 0; c < NCM; ++c) {
      if (c < NC) {
        uint32_t mk, mis, chk8 = 0;
        switch (c) {   // constant bit positions
#define STORM_ROW0(CC)                                                                          \
  case CC:                                                                                      \
    mk = pbit<CC>(Pm);                                                                          \
    mis = pbit<CC>(MIS);                                                                        \
    if (COLOR) chk8 = pmove<CC, 8>(X0) | pmove<CC, 9>(X1);                                      \
    break;
          STORM_ROW0(0) STORM_ROW0(1) STORM_ROW0(2) STORM_ROW0(3) STORM_ROW0(4) STORM_ROW0(5) STORM_ROW0(6)
          STORM_ROW0(7) STORM_ROW0(8) STORM_ROW0(9) STORM_ROW0(10) STORM_ROW0(11) STORM_ROW0(12) STORM_ROW0(13)
          STORM_ROW0(14)
#undef STORM_ROW0
          default: mk = mis = 0; break;
        }
        sc[c] = psel(mk, MIN2, psel(mis, s_x, s_m));
        if (COLOR) {   // chk = auto = r0 ^ g, fresh = q0 on a mismatch (masked: all 0)
          const uint32_t cchk8 = chk8 & ~mk;
          meta[c] = pk(3) | cchk8 | (cchk8 << 2) | (q12 & mis & ~mk);
        } else {
          meta[c] = pk(3);
        }
      } else {
        sc[c] = MIN2;
        meta[c] = BORDER2;
      }
    }
  }
  if (TB) {   // row 0: every cell from the corner (3)
#pragma unroll
    for (int w = 0; w < DW; ++w) { tbw[128 * w] = 0x33333333u; tbw[128 * w + 64] = 0x33333333u; }
  }
  const uint32_t L2 = (uint32_t)(uint16_t)A.L | ((uint32_t)(uint16_t)B.L << 16);
  const uint32_t top2 = pk(k.top_match), t2 = pk(k.t);
  uint32_t bp = ((uint32_t)(uint16_t)((COLOR ? A.L - 2 : A.L) * k.top_match)) |
                ((uint32_t)(uint16_t)((COLOR ? B.L - 2 : B.L) * k.top_match) << 16);
  uint32_t S = 0x7fff7fffu, score = MIN2, done = 0, bestc = 0xffffffffu;
  const int maxL = A.L > B.L ? A.L : B.L;
  for (int i = 1; i < maxL && done != 0xffffffffu; ++i) {
    A.advance(g, i, has_qual);
    B.advance(g, i, has_qual);
    const uint32_t ri = A.rcode(i) | (B.rcode(i) << 16), qi = A.qlevel(i) | (B.qlevel(i) << 16);
    const uint32_t qsel = (qi << 1) | (qi << 9) | 0x01000100u;
    const uint32_t s_m = __builtin_amdgcn_perm(tm23, tm01, qsel), s_x = __builtin_amdgcn_perm(tx23, tx01, qsel);
    const uint32_t Plo = __builtin_amdgcn_perm(B.wlo(), A.wlo(), 0x05040100u);
    const uint32_t Phi = __builtin_amdgcn_perm(B.whi(), A.whi(), 0x05040100u);
    const uint32_t Pm = __builtin_amdgcn_perm(B.wmk(), A.wmk(), 0x05040100u);
    const uint32_t X0 = Plo ^ psub16(0, ri & 0x00010001u), X1 = Phi ^ psub16(0, (ri >> 1) & 0x00010001u);
    const uint32_t MIS = X0 | X1;
    const uint32_t ri8 = ri << 8, q12 = qi << 12;
    uint32_t mol = MIN2;
    uint32_t left_sc = MIN2, left_meta = BORDER2;
    uint32_t dA[DW], dB[DW];
#pragma unroll
    for (int w = 0; w < DW; ++w) dA[w] = dB[w] = 0;
#pragma unroll
    for (int c = 0; c < NCM; ++c) {
      if (c < NC) {
        uint32_t mk, mis, chk8 = 0;
        switch (c) {
#define STORM_ROWC(CC)                                                                          \
  case CC:                                                                                      \
    mk = pbit<CC>(Pm);                                                                          \
    mis = pbit<CC>(MIS);                                                                        \
    if (COLOR) chk8 = pmove<CC, 8>(X0) | pmove<CC, 9>(X1);                                      \
    break;
          STORM_ROWC(0) STORM_ROWC(1) STORM_ROWC(2) STORM_ROWC(3) STORM_ROWC(4) STORM_ROWC(5) STORM_ROWC(6)
          STORM_ROWC(7) STORM_ROWC(8) STORM_ROWC(9) STORM_ROWC(10) STORM_ROWC(11) STORM_ROWC(12) STORM_ROWC(13)
          STORM_ROWC(14)
#undef STORM_ROWC
          default: mk = mis = 0; break;
        }
        const uint32_t cs = psel(mk, MIN2, psel(mis, s_x, s_m));
        const uint32_t csc = sc[c], cmeta = meta[c];
        const uint32_t tsc = (c + 1 < NCM) ? sc[(c + 1) < NCM ? c + 1 : c] : MIN2;
        const uint32_t tmeta = (c + 1 < NCM) ? meta[(c + 1) < NCM ? c + 1 : c] : BORDER2;
        uint32_t ms = cs, cchk8 = 0, xw = 0, iauto = 0, ichk0 = 0, xw0 = 0, xauto = 0, mis_nm = 0;
        if (COLOR) {
          cchk8 = chk8 & ~mk;
          mis_nm = mis & ~mk;   // the cell's checksum is not zero
          xw = (cmeta ^ cchk8) & 0x03000300u;
          iauto = cmeta & 0x0C000C00u;
          ichk0 = m_zero(cmeta & 0x03000300u);
          xw0 = m_zero(xw);
          xauto = m_zero((xw << 2) ^ iauto);
          // alignment.c:465-468: a mismatch that corrects the pending checksum is refunded
          const uint32_t refund = ~ichk0 & (xw0 | xauto) & ~mk & pu(ps(cs) >> (s2){15, 15});
          const uint32_t neg = psub16(0, cs);
          ms = psel(refund, pu(ps(neg) + ps(neg)), cs);
        }
        const uint32_t diag = padd(csc, ms);
        const uint32_t t_lo8 = COLOR ? (tmeta & 0x00ff00ffu) : tmeta;
        const uint32_t l_lo8 = COLOR ? (left_meta & 0x00ff00ffu) : left_meta;
        // a cell's `from` is 1 (left), 2 (top) or 3 (corner; also row 0 and the band border): a gap
        // extends from the top cell iff its from is 2 (bit 0 clear), from the left iff 1 (bit 1 clear)
        const uint32_t t_try = padd(tsc, pmad(t_lo8 & 0x00010001u, DGAP2, EXT2));
        const uint32_t take_t = m_gt(I4_2, t_lo8) & m_gt(t_try, diag);
        const uint32_t after_t = psel(take_t, t_try, diag);
        const uint32_t l_try = padd(left_sc, pmad((l_lo8 >> 1) & 0x00010001u, DGAP2, EXT2));
        const uint32_t take_l = m_gt(I4_2, l_lo8) & m_gt(l_try, after_t);
        const uint32_t nsc = psel(take_l, l_try, after_t);
        const uint32_t d_lo = (COLOR ? (cmeta & 0x00ff00ffu) : cmeta) | 0x00030003u;
        const uint32_t t_lo = (t_lo8 & 0x00fc00fcu) + 0x00060006u;
        const uint32_t l_lo = (l_lo8 & 0x00fc00fcu) + 0x00050005u;
        const uint32_t n_lo = psel(take_l, l_lo, psel(take_t, t_lo, d_lo));
        uint32_t nmeta = n_lo;
        if (COLOR) {
          const uint32_t fw = cmeta & 0x70007000u;
          const uint32_t f1 = m_gt(fw, pk(0x1000));
          const uint32_t xwK = xw | iauto | psub16(fw, pk(0x1000));
          const uint32_t A1w = xw | (xw << 2) | q12;
          const uint32_t a2 = ~ichk0 & xauto, a3 = ~ichk0 & ~a2 & xw0;
          const uint32_t m_a = psel(ichk0, A1w, psel(a2, 0u, psel(f1, xwK, psel(a3, 0u, A1w))));
          const uint32_t b1 = ~m_zero(iauto), b_keep = b1 & f1 & ~xw0;
          const uint32_t m_b = psel(b1, psel(b_keep, xwK, xw), 0u);
          const uint32_t d_hi = psel(mis_nm, m_a, m_b);
          const uint32_t c_af = (cchk8 << 2) | (q12 & mis_nm);
          const uint32_t tfw = tmeta & 0x70007000u, lfw = left_meta & 0x70007000u;
          const uint32_t gj8 = (chk8 ^ ri8) & 0x03000300u;
          const uint32_t t_hi = ((tmeta ^ ri8) & 0x03000300u) |
                                psel(m_gt(tfw, pk(0x1000)), iauto | psub16(tfw, pk(0x1000)), c_af);
          const uint32_t l_hi = ((left_meta ^ gj8) & 0x03000300u) |
                                psel(m_gt(lfw, pk(0x1000)), iauto | psub16(lfw, pk(0x1000)), c_af);
          nmeta = n_lo | psel(take_l, l_hi, psel(take_t, t_hi, d_hi));
        }
        if (TB) {   // direction nibble: from | 4 when the score rose from the corner
          const uint32_t nib = (n_lo & 0x00030003u) | (m_gt(nsc, csc) & 0x00040004u);
          dA[c >> 3] |= (nib & 0xfu) << ((c & 7) * 4);
          dB[c >> 3] |= ((nib >> 16) & 0xfu) << ((c & 7) * 4);
        }
        sc[c] = nsc;
        meta[c] = nmeta;
        left_sc = nsc;
        left_meta = nmeta;
        mol = pmax(mol, nsc);
      }
    }
    if (TB) {
#pragma unroll
      for (int w = 0; w < DW; ++w) { tbw[128 * (i * DW + w)] = dA[w]; tbw[128 * (i * DW + w) + 64] = dB[w]; }
    }
    const uint32_t ip = pk(i);
    const uint32_t act = m_gt(L2, ip);                    // row i < L
    S = psel(act, pmin(S, padd(mol, bp)), S);
    bp = psub16(bp, top2);
    const uint32_t last = m_zero(psub16(L2, pk(i + 1)));  // row i = L - 1
    score = psel(last, mol, score);
    if (TB && last) {   // the leftmost column holding the last row's maximum (alignment.c:621-630)
      uint32_t bc = 0xffffffffu;
#pragma unroll
      for (int c = NCM - 1; c >= 0; --c)
        if (c < NC) bc = psel(m_zero(sc[c] ^ mol), pk(c), bc);
      bestc = psel(last, bc, bestc);
    }
    done |= ~act | last | m_gt(t2, S);
  }
  const int SA = (int)(int16_t)(S & 0xffffu), SB = (int)(int16_t)(S >> 16);
  const int scA = (int)(int16_t)(score & 0xffffu), scB = (int)(int16_t)(score >> 16);
  outA.score = SA < k.t ? MIN_SCORE : scA;
  outA.S = SA;
  outA.win = A.win;
  outB.score = SB < k.t ? MIN_SCORE : scB;
  outB.S = SB;
  outB.win = B.win;
  outA.best_j = outA.best_i0 = outA.best_j0 = -1;
  outB.best_j = outB.best_i0 = outB.best_j0 = -1;
  outA.tb_len = outB.tb_len = 0;
  if (TB) {
    const int bcA = (int)(int16_t)(bestc & 0xffffu), bcB = (int)(int16_t)(bestc >> 16);
    if (bcA >= 0 && outA.score > k.t && SA >= k.t) {
      outA.best_j = (int16_t)(A.L - 1 + bcA);
      if (tbA) dp_walk<IM, COLOR>(g, A.pl, A.L, A.win, NC, bcA, tbw, tbA, outA);
    }
    if (bcB >= 0 && outB.score > k.t && SB >= k.t) {
      outB.best_j = (int16_t)(B.L - 1 + bcB);
      if (tbB) dp_walk<IM, COLOR>(g, B.pl, B.L, B.win, NC, bcB, tbw + 64, tbB, outB);
    }
  }
}

}  // namespace storm
