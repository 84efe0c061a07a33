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

// One candidate: read strand rs (2*read + strand), diagonal h.  tbm is this thread's scratch of
// L*(2I+1) direction bytes; the traceback (alignment order, 2 codes per byte) goes to tb.
template <int IM, bool COLOR>
__device__ AlignRec dp_one(const GenomeDev& g, const ReadsDev& r, const ScoreDev& scd, int rs, int h, uint8_t* tbm,
                           uint8_t* tb) {
  constexpr int NCM = 2 * IM + 1;
  const int L = r.len[rs >> 1];
  const int I = scd.I, d = scd.d;
  const int NC = 2 * I + 1;
  const int win = (h - I > -d) ? h - I : -d;
  const uint32_t* pl = read_planes(r, rs);
  const int top_match = scd.ps[0][3];

  int sc[NCM];
  int meta[NCM];
  // window codes of a row as bit-plane words: bit c of (w*lo/hi/mask) is genome code win + i + c
  const int64_t wbase = (int64_t)win + GPAD;
  // row 0 (alignment.c:326-360)
  {
    const int r0 = rcode(pl, 0), q0 = rqual(r, rs, 0);
    const uint32_t wl0 = gbit_word(g.lo, wbase), wh0 = gbit_word(g.hi, wbase), wm0 = gbit_word(g.mask, wbase);
    const uint32_t wl1 = NCM > 32 ? gbit_word(g.lo, wbase + 32) : 0u, wh1 = NCM > 32 ? gbit_word(g.hi, wbase + 32) : 0u;
    const uint32_t wm1 = NCM > 32 ? gbit_word(g.mask, wbase + 32) : 0u;
#pragma unroll
    for (int c = 0; c < NCM; ++c) {
      if (c < NC) {
        const uint32_t L0 = c < 32 ? wl0 : wl1, H0 = c < 32 ? wh0 : wh1, K0 = c < 32 ? wm0 : wm1;
        const int gj = (int)(((L0 >> (c & 31)) & 1u) | (((H0 >> (c & 31)) & 1u) << 1));
        if (((K0 >> (c & 31)) & 1u) == 0) {
          if (COLOR) {
            const int chk = r0 ^ gj;
            sc[c] = scd.ps[chk ? 1 : 0][q0];
            meta[c] = m_make(3, 0, chk, chk, chk ? q0 : 0);
          } else {
            sc[c] = scd.ps[r0 != gj ? 1 : 0][q0];
            meta[c] = m_make(3, 0, 0, 0, 0);
          }
        } else {
          sc[c] = MIN_SCORE;
          meta[c] = m_make(3, 0, 0, 0, 0);
        }
        tbm[c] = 3;
      }
    }
  }
  int bp = (COLOR ? (L - 2) : L) * top_match;
  int S = 0x7fffffff;
  for (int i = 1; i < L; ++i) {
    const int ri = rcode(pl, i), qi = rqual(r, rs, i);
    const uint32_t wl0 = gbit_word(g.lo, wbase + i), wh0 = gbit_word(g.hi, wbase + i), wm0 = gbit_word(g.mask, wbase + i);
    const uint32_t wl1 = NCM > 32 ? gbit_word(g.lo, wbase + i + 32) : 0u, wh1 = NCM > 32 ? gbit_word(g.hi, wbase + i + 32) : 0u;
    const uint32_t wm1 = NCM > 32 ? gbit_word(g.mask, wbase + i + 32) : 0u;
    int max_on_line = MIN_SCORE;
    int left_sc = MIN_SCORE, left_meta = m_make(3, I, 0, 0, 0);   // border (alignment.c:303-318)
#pragma unroll
    for (int c = 0; c < NCM; ++c) {
      if (c < NC) {
        const uint32_t Lw = c < 32 ? wl0 : wl1, Hw = c < 32 ? wh0 : wh1, Kw = c < 32 ? wm0 : wm1;
        const int gj = (int)(((Lw >> (c & 31)) & 1u) | (((Hw >> (c & 31)) & 1u) << 1));
        const bool masked = ((Kw >> (c & 31)) & 1u) != 0;
        // this cell before the recurrence (init_matrix)
        int cs, cchk = 0, cauto = 0, cfresh = 0;
        if (!masked) {
          if (COLOR) {
            cchk = ri ^ gj; cauto = cchk; cfresh = cchk ? qi : 0;
            cs = scd.ps[cchk ? 1 : 0][qi];
          } else {
            cs = scd.ps[ri != gj ? 1 : 0][qi];
          }
        } else {
          cs = MIN_SCORE;
        }
        const int csc = sc[c], cmeta = meta[c];                          // corner (i-1, j-1)
        const int tsc = (c + 1 < NC) ? sc[(c + 1) < NCM ? c + 1 : c] : MIN_SCORE;   // top (i-1, j)
        const int tmeta = (c + 1 < NC) ? meta[(c + 1) < NCM ? c + 1 : c] : m_make(3, I, 0, 0, 0);
        int match_score = cs;
        if (COLOR) {   // alignment.c:465-468
          const int prev = m_chk(cmeta), ref = m_auto(cmeta);
          if (prev && (((prev ^ cchk) == 0) || ((prev ^ cchk) == ref)) && cs != MIN_SCORE && cs < 0) match_score = -2 * cs;
        }
        int nsc = csc + match_score;
        int it = 0;   // 0 corner, 1 top, 2 left
        if (m_ind(tmeta) < I) {
          const int tmp = tsc + ((m_from(tmeta) ^ 2) ? scd.gap_open : scd.gap_ext);
          if (tmp > nsc) { it = 1; nsc = tmp; }
        }
        if (m_ind(left_meta) < I) {
          const int tmp = left_sc + ((m_from(left_meta) ^ 1) ? scd.gap_open : scd.gap_ext);
          if (tmp > nsc) { it = 2; nsc = tmp; }
        }
        int nfrom, nind, nchk = cchk, nauto = cauto, nfresh = cfresh;
        if (it == 0) {
          nfrom = 3;
          nind = m_ind(cmeta);
          if (COLOR) {   // alignment.c:533-578
            const int ichk = m_chk(cmeta), iauto = m_auto(cmeta), ifresh = m_fresh(cmeta);
            nchk = cchk ^ ichk;
            if (ichk != nchk) {
              if (ichk == 0) { nauto = nchk; nfresh = qi; }
              else if (nchk == iauto) { nauto = 0; nchk = 0; nfresh = 0; }
              else if (nchk == 0) {
                if (ifresh - 1 > 0) { nauto = iauto; nfresh = ifresh - 1; } else { nauto = 0; nfresh = 0; }
              } else {
                if (ifresh - 1 > 0) { nauto = iauto; nfresh = ifresh - 1; } else { nauto = nchk; nfresh = qi; }
              }
            } else {
              if (iauto != 0) {
                if (ifresh - 1 > 0) {
                  if (nchk == 0) { nauto = 0; nfresh = 0; } else { nauto = iauto; nfresh = ifresh - 1; }
                }
              } else {
                nchk = 0; nauto = 0; nfresh = 0;
              }
            }
          }
        } else if (it == 1) {   // alignment.c:580-589
          nfrom = 2;
          nind = m_ind(tmeta) + 1;
          if (COLOR) {
            nchk = m_chk(tmeta) ^ ri;
            if (m_fresh(tmeta) - 1 > 0) { nfresh = m_fresh(tmeta) - 1; nauto = m_auto(cmeta); }
          }
        } else {   // alignment.c:590-599
          nfrom = 1;
          nind = m_ind(left_meta) + 1;
          if (COLOR) {
            nchk = m_chk(left_meta) ^ gj;
            if (m_fresh(left_meta) - 1 > 0) { nfresh = m_fresh(left_meta) - 1; nauto = m_auto(cmeta); }
          }
        }
        tbm[i * NC + c] = (uint8_t)(nfrom | ((nsc > csc) ? 4 : 0));
        sc[c] = nsc;
        meta[c] = m_make(nfrom, nind, nchk, nauto, nfresh);
        left_sc = nsc;
        left_meta = meta[c];
        if (max_on_line < nsc) max_on_line = nsc;
      }
    }
    if (max_on_line + bp < S) S = max_on_line + bp;
    bp -= top_match;
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
  if (bc >= 0 && best > scd.t && S >= scd.t) {
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
