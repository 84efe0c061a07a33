// dev_align.h -- one banded DP + traceback (used by k_align and k_fused); semantics in k_align.hip.
#pragma once
#include "storm_kernels.h"

namespace storm {

// meta word of a cell: from (2 bits) | indels << 2 (6 bits) | chk << 8 | autoc << 10 | fresh << 12
__device__ __forceinline__ int m_from(int m) { return m & 3; }
__device__ __forceinline__ int m_ind(int m) { return (m >> 2) & 63; }
__device__ __forceinline__ int m_chk(int m) { return (m >> 8) & 3; }
__device__ __forceinline__ int m_auto(int m) { return (m >> 10) & 3; }
__device__ __forceinline__ int m_fresh(int m) { return (m >> 12) & 7; }
__device__ __forceinline__ int m_make(int from, int ind, int chk, int autoc, int fresh) {
  return from | (ind << 2) | (chk << 8) | (autoc << 10) | (fresh << 12);
}

// a ? x : y as mask arithmetic: the compiler keeps it a straight-line select instead of a
// divergent branch around the computation of the unused operand
__device__ __forceinline__ int msel(bool a, int x, int y) { return y ^ ((y ^ x) & -(int)a); }

// One candidate: read strand rs (2*read + strand), diagonal h.
// TB = true: tbm is this thread's scratch of L*(2I+1) direction bytes and, when the record can be
// accepted, the traceback (alignment order, 2 codes per byte) goes to tb.
// TB = false: the score pass alone, nothing is stored.  A record can only ever be accepted when
// score > T and S >= T with T >= t (T starts at -t and only rises, run.c:568,683-685), so this
// pass stops at the first row where the bound drops below t (S < t then holds) and returns
// score = MIN_SCORE; otherwise it returns the same score, S and best_j as the TB pass.
template <int IM, bool COLOR, bool TB = true>
__device__ AlignRec dp_one(const GenomeDev& g, const ReadsDev& r, const ScoreDev& scd, int rs, int h, uint8_t* tbm,
                           uint8_t* tb) {
  constexpr int NCM = 2 * IM + 1;
  const int L = r.len[rs >> 1];
  const int I = scd.I, d = scd.d;
  const int NC = 2 * I + 1;
  const int win = (h - I > -d) ? h - I : -d;
  const uint32_t* pl = read_planes(r, rs);
  const int top_match = scd.ps[0][3];
  const int gap_open = scd.gap_open, gap_ext = scd.gap_ext;
  const int I4 = 4 * scd.I;   // a meta low byte below 4I has fewer than I indels
  // pair scores by quality level, in registers: a lane-indexed read of the argument block
  // would be a memory round trip per cell
  const int pm0 = scd.ps[0][0], pm1 = scd.ps[0][1], pm2 = scd.ps[0][2], pm3 = scd.ps[0][3];
  const int px0 = scd.ps[1][0], px1 = scd.ps[1][1], px2 = scd.ps[1][2], px3 = scd.ps[1][3];
  auto level_score = [](int q, int s0, int s1, int s2, int s3) {
    return msel(q & 2, msel(q & 1, s3, s2), msel(q & 1, s1, s0));
  };

  int sc[NCM];
  int meta[NCM];
  // Sliding bit-plane windows of the genome: bit c of row i's word is code win + i + c.  Per plane
  // NW words from word index wi are held in registers; a new word is fetched every 32 rows, one
  // word ahead of its use, so the row loop issues no dependent loads.
  constexpr int NW = NCM > 32 ? 4 : 3;
  const int64_t wbase = (int64_t)win + GPAD;
  int64_t wi = wbase >> 5;
  int sh = (int)(wbase & 31);
  uint32_t gl[NW], gh[NW], gm[NW];
#pragma unroll
  for (int k = 0; k < NW; ++k) { gl[k] = g.lo[wi + k]; gh[k] = g.hi[wi + k]; gm[k] = g.mask[wi + k]; }
  // read codes and quality levels, one word per 32 (codes) or 16 (levels) rows
  const uint32_t* qp = r.qual + (int64_t)rs * QW;
  uint32_t rlo = pl[0], rhi = pl[RW], rq = r.has_qual ? qp[0] : 0xffffffffu;
  // row 0 (alignment.c:326-360)
  {
    const int r0 = (int)((rlo & 1u) | ((rhi & 1u) << 1)), q0 = (int)(rq & 3u);
    const int s_m = level_score(q0, pm0, pm1, pm2, pm3), s_x = level_score(q0, px0, px1, px2, px3);
    const uint32_t wl0 = __builtin_amdgcn_alignbit(gl[1], gl[0], sh), wh0 = __builtin_amdgcn_alignbit(gh[1], gh[0], sh);
    const uint32_t wm0 = __builtin_amdgcn_alignbit(gm[1], gm[0], sh);
    const uint32_t wl1 = NCM > 32 ? __builtin_amdgcn_alignbit(gl[2], gl[1], sh) : 0u;
    const uint32_t wh1 = NCM > 32 ? __builtin_amdgcn_alignbit(gh[2], gh[1], sh) : 0u;
    const uint32_t wm1 = NCM > 32 ? __builtin_amdgcn_alignbit(gm[2], gm[1], sh) : 0u;
#pragma unroll
    for (int c = 0; c < NCM; ++c) {
      if (c < NC) {
        const uint32_t L0 = c < 32 ? wl0 : wl1, H0 = c < 32 ? wh0 : wh1, K0 = c < 32 ? wm0 : wm1;
        const int gj = (int)(((L0 >> (c & 31)) & 1u) | (((H0 >> (c & 31)) & 1u) << 1));
        if (((K0 >> (c & 31)) & 1u) == 0) {
          if (COLOR) {
            const int chk = r0 ^ gj;
            sc[c] = chk ? s_x : s_m;
            meta[c] = m_make(3, 0, chk, chk, chk ? q0 : 0);
          } else {
            sc[c] = (r0 != gj) ? s_x : s_m;
            meta[c] = m_make(3, 0, 0, 0, 0);
          }
        } else {
          sc[c] = MIN_SCORE;
          meta[c] = m_make(3, 0, 0, 0, 0);
        }
        if (TB) tbm[c] = 3;
      } else {   // columns past the band hold the border, so the top neighbour needs no test
        sc[c] = MIN_SCORE;
        meta[c] = m_make(3, I, 0, 0, 0);
      }
    }
  }
  int bp = (COLOR ? (L - 2) : L) * top_match;
  int S = 0x7fffffff;
  for (int i = 1; i < L; ++i) {
    // next row: advance the genome windows by one code
    if (++sh == 32) {
      sh = 0;
      ++wi;
#pragma unroll
      for (int k = 0; k + 1 < NW; ++k) { gl[k] = gl[k + 1]; gh[k] = gh[k + 1]; gm[k] = gm[k + 1]; }
      gl[NW - 1] = g.lo[wi + NW - 1];
      gh[NW - 1] = g.hi[wi + NW - 1];
      gm[NW - 1] = g.mask[wi + NW - 1];
    }
    if ((i & 31) == 0) { rlo = pl[i >> 5]; rhi = pl[RW + (i >> 5)]; }
    if ((i & 15) == 0 && r.has_qual) rq = qp[i >> 4];
    const int ri = (int)(((rlo >> (i & 31)) & 1u) | (((rhi >> (i & 31)) & 1u) << 1));
    const int qi = (int)((rq >> ((i & 15) << 1)) & 3u);
    const int s_m = level_score(qi, pm0, pm1, pm2, pm3), s_x = level_score(qi, px0, px1, px2, px3);
    const uint32_t wl0 = __builtin_amdgcn_alignbit(gl[1], gl[0], sh), wh0 = __builtin_amdgcn_alignbit(gh[1], gh[0], sh);
    const uint32_t wm0 = __builtin_amdgcn_alignbit(gm[1], gm[0], sh);
    const uint32_t wl1 = NCM > 32 ? __builtin_amdgcn_alignbit(gl[2], gl[1], sh) : 0u;
    const uint32_t wh1 = NCM > 32 ? __builtin_amdgcn_alignbit(gh[2], gh[1], sh) : 0u;
    const uint32_t wm1 = NCM > 32 ? __builtin_amdgcn_alignbit(gm[2], gm[1], sh) : 0u;
    // bit c set when read code ri differs from genome code win + i + c
    const uint32_t rb0 = 0u - (uint32_t)(ri & 1), rb1 = 0u - (uint32_t)((ri >> 1) & 1);
    const uint32_t mis0 = (wl0 ^ rb0) | (wh0 ^ rb1), mis1 = NCM > 32 ? ((wl1 ^ rb0) | (wh1 ^ rb1)) : 0u;
    int max_on_line = MIN_SCORE;
    int left_sc = MIN_SCORE, left_meta = m_make(3, I, 0, 0, 0);   // border (alignment.c:303-318)
#pragma unroll
    for (int c = 0; c < NCM; ++c) {
      if (c < NC) {
        const uint32_t Lw = c < 32 ? wl0 : wl1, Hw = c < 32 ? wh0 : wh1, Kw = c < 32 ? wm0 : wm1;
        const int gj = (int)(((Lw >> (c & 31)) & 1u) | (((Hw >> (c & 31)) & 1u) << 1));
        const bool masked = ((Kw >> (c & 31)) & 1u) != 0;
        // this cell before the recurrence (init_matrix): pair score, or MIN_SCORE when masked;
        // in color mode the checksum, auto-correction and freshness of a mismatch
        const bool mis = (((c < 32 ? mis0 : mis1) >> (c & 31)) & 1u) != 0;
        const int cchk = msel(masked, 0, ri ^ gj);
        const int cs = msel(masked, MIN_SCORE, msel(mis, s_x, s_m));
        const int cauto = cchk, cfresh = msel(cchk != 0, qi, 0);
        const int csc = sc[c], cmeta = meta[c];                          // corner (i-1, j-1)
        const int tsc = (c + 1 < NCM) ? sc[(c + 1) < NCM ? c + 1 : c] : MIN_SCORE;   // top (i-1, j)
        const int tmeta = (c + 1 < NCM) ? meta[(c + 1) < NCM ? c + 1 : c] : m_make(3, I, 0, 0, 0);
        // The recurrence in select form (alignment.c:465-604): every lane runs the same
        // instructions whatever its record, with no divergent branches per cell.
        int match_score = cs;
        if (COLOR) {   // alignment.c:465-468
          const int prev = m_chk(cmeta), ref = m_auto(cmeta);
          // non-short-circuit & and |: short-circuit operators become divergent branches
          const bool refund = (prev != 0) & (((prev ^ cchk) == 0) | ((prev ^ cchk) == ref)) & (cs != MIN_SCORE) & (cs < 0);
          match_score = msel(refund, -2 * cs, cs);
        }
        const int diag = csc + match_score;
        const int t_lo8 = COLOR ? (tmeta & 0xff) : tmeta, l_lo8 = COLOR ? (left_meta & 0xff) : left_meta;
        const int t_try = tsc + msel((t_lo8 & 3) == 2, gap_ext, gap_open);
        const bool take_t = (t_lo8 < I4) & (t_try > diag);   // indels < I

        const int after_t = msel(take_t, t_try, diag);
        const int l_try = left_sc + msel((l_lo8 & 3) == 1, gap_ext, gap_open);
        const bool take_l = (l_lo8 < I4) & (l_try > after_t);
        const int nsc = msel(take_l, l_try, after_t);
        // low byte of the new meta word (from | indels << 2): the corner's indels with from = 3, or a
        // neighbour's indels + 1 with from = 2 (top) / 1 (left); indels <= 31, so +4 never carries out
        const int d_lo = (COLOR ? (cmeta & 0xff) : cmeta) | 3;
        const int t_lo = (t_lo8 & 0xfc) + 6;
        const int l_lo = (l_lo8 & 0xfc) + 5;
        const int n_lo = msel(take_l, l_lo, msel(take_t, t_lo, d_lo));
        const int nfrom = n_lo & 3;
        int nchk = 0, nauto = 0, nfresh = 0;
        if (COLOR) {
          // diagonal move (alignment.c:533-578)
          const int ichk = m_chk(cmeta), iauto = m_auto(cmeta), ifresh = m_fresh(cmeta);
          const int x = cchk ^ ichk;
          const bool f1 = ifresh > 1;
          // a colour mismatch here (cchk != 0): the four cases of the reference, in order
          const bool a1 = ichk == 0, a2 = (!a1) & (x == iauto), a3 = (!a1) & (!a2) & (x == 0);
          const int a_chk = msel(a2, 0, x);
          const int a_auto = msel(a1, x, msel(a2, 0, msel(f1, iauto, msel(a3, 0, x))));
          const int a_fresh = msel(a1, qi, msel(a2, 0, msel(f1, ifresh - 1, msel(a3, 0, qi))));
          // a colour match (cchk == 0)
          const bool b1 = iauto != 0, b_keep = b1 & f1 & (x != 0);
          const int b_chk = msel(b1, x, 0);
          const int b_auto = msel(b_keep, iauto, 0);
          const int b_fresh = msel(b_keep, ifresh - 1, 0);
          const bool cmis = cchk != 0;
          const int d_chk = msel(cmis, a_chk, b_chk), d_auto = msel(cmis, a_auto, b_auto);
          const int d_fresh = msel(cmis, a_fresh, b_fresh);
          // gap from the top (alignment.c:580-589) or from the left (alignment.c:590-599)
          const int tf = m_fresh(tmeta), lf = m_fresh(left_meta);
          const int t_chk = m_chk(tmeta) ^ ri, l_chk = m_chk(left_meta) ^ gj;
          const int t_auto = msel(tf > 1, m_auto(cmeta), cauto), l_auto = msel(lf > 1, m_auto(cmeta), cauto);
          const int t_fresh = msel(tf > 1, tf - 1, cfresh), l_fresh = msel(lf > 1, lf - 1, cfresh);
          nchk = msel(take_l, l_chk, msel(take_t, t_chk, d_chk));
          nauto = msel(take_l, l_auto, msel(take_t, t_auto, d_auto));
          nfresh = msel(take_l, l_fresh, msel(take_t, t_fresh, d_fresh));
        }
        if (TB) tbm[i * NC + c] = (uint8_t)(nfrom | ((nsc > csc) ? 4 : 0));
        sc[c] = nsc;
        meta[c] = COLOR ? (n_lo | (nchk << 8) | (nauto << 10) | (nfresh << 12)) : n_lo;
        left_sc = nsc;
        left_meta = meta[c];
        if (max_on_line < nsc) max_on_line = nsc;
      }
    }
    if (max_on_line + bp < S) S = max_on_line + bp;
    bp -= top_match;
    if (!TB && S < scd.t) {   // can never be accepted: the rest of the rows cannot raise S
      AlignRec o;
      o.score = MIN_SCORE;
      o.S = S;
      o.win = win;
      o.best_j = -1;
      o.best_i0 = -1;
      o.best_j0 = -1;
      o.tb_len = 0;
      return o;
    }
  }
  // leftmost maximum of the last row (alignment.c:621-630)
  int best = MIN_SCORE, bc = -1;
#pragma unroll
  for (int c = 0; c < NCM; ++c)
    if (c < NC && best < sc[c]) { best = sc[c]; bc = c; }

  AlignRec o;
  o.score = best;
  o.S = S;
  o.win = win;
  o.best_j = (int16_t)(bc >= 0 ? L - 1 + bc : -1);
  o.best_i0 = -1;
  o.best_j0 = -1;
  o.tb_len = 0;
  // A record can be accepted only if score > T and S >= T with T >= t (run.c:568,675-685).
  if (TB && bc >= 0 && best > scd.t && S >= scd.t) {
    int i = L - 1, c = bc, n = 0, color = 0;
    while (i >= 0 && c >= 0 && c < NC) {
      const int j = i + c;
      o.best_i0 = (int16_t)i;
      o.best_j0 = (int16_t)j;
      const int cell = tbm[i * NC + c];
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
  return o;
}

}  // namespace storm
