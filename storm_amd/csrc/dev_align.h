// dev_align.h -- banded glocal DP + traceback (used by k_align and k_fused); semantics in k_align.hip.
#pragma once
#include <type_traits>
#include "storm_kernels.h"

namespace storm {

// meta word of a cell: from (2 bits) | indels << 2 (6 bits) | chk << 8 | autoc << 10 | fresh << 12,
// and in bits 16-31 the signed gap penalty of a vertical gap step out of the cell (extend when the
// cell was reached from the top, else open: alignment.c:386-388), so the cell below adds it as is
__device__ __forceinline__ int m_from(int m) { return m & 3; }
__device__ __forceinline__ int m_ind(int m) { return (m >> 2) & 63; }
__device__ __forceinline__ int m_chk(int m) { return (m >> 8) & 3; }
__device__ __forceinline__ int m_auto(int m) { return (m >> 10) & 3; }
__device__ __forceinline__ int m_fresh(int m) { return (m >> 12) & 7; }
__device__ __forceinline__ int m_make(int from, int ind, int chk, int autoc, int fresh) {
  return from | (ind << 2) | (chk << 8) | (autoc << 10) | (fresh << 12);
}

// a ? x : y over operands that are both computed: one v_cndmask on the compare's lane mask (no
// branch; mask arithmetic would first turn the mask into a 0 / -1 register)
__device__ __forceinline__ int msel(bool a, int x, int y) { return a ? x : y; }

// Scoring constants of the DP in registers (a lane-indexed read of the argument block would be a
// memory round trip per cell).
struct DPConst {
  int I, d, NC, I4, t, top_match, gap_open, gap_ext;
  int m_diag, m_top, m_left;   // the new meta word's constant part: from, indels step, vertical gap penalty
  int pm0, pm1, pm2, pm3, px0, px1, px2, px3;   // pair_scores[match / mismatch][level]
  __device__ explicit DPConst(const ScoreDev& s)
      : I(s.I), d(s.d), NC(2 * s.I + 1), I4(4 * s.I), t(s.t), top_match(s.ps[0][3]), gap_open(s.gap_open),
        gap_ext(s.gap_ext), pm0(s.ps[0][0]), pm1(s.ps[0][1]), pm2(s.ps[0][2]), pm3(s.ps[0][3]), px0(s.ps[1][0]),
        px1(s.ps[1][1]), px2(s.ps[1][2]), px3(s.ps[1][3]) {
    const int vo = (int)((uint32_t)(uint16_t)(short)s.gap_open << 16), ve = (int)((uint32_t)(uint16_t)(short)s.gap_ext << 16);
    m_diag = 3 | vo;       // from the corner: indels kept, a vertical step out opens a gap
    m_top = 6 | ve;        // from the top: indels + 1, a vertical step out extends the gap
    m_left = 5 | vo;       // from the left: indels + 1, a vertical step out opens a gap
  }
  __device__ int match(int q) const { return msel(q & 2, msel(q & 1, pm3, pm2), msel(q & 1, pm1, pm0)); }
  __device__ int mismatch(int q) const { return msel(q & 2, msel(q & 1, px3, px2), msel(q & 1, px1, px0)); }
};

// Color space: the diagonal move's color state as a table.  A cell's state is its checksum,
// auto-correction and mismatch freshness (2 bits each, freshness <= 3: it starts at a quality
// level and only decreases), packed as chk | auto << 2 | fresh << 4 (the cell's meta bits 8-13).
// Entry [corner state | chk0 << 6 | q << 8] -- chk0 the cell's own read ^ genome color (0 for a
// masked cell), q its quality level -- holds the state a diagonal move from that corner gives the
// cell (alignment.c:533-578) and, at bit 6, whether the corner's state turns the cell's mismatch
// penalty into a refund (alignment.c:465-468: pending checksum, and the composed checksum is the
// identity or the corner's auto-correction; the caller adds "penalty < 0, cell not masked").
// The table is built at compile time; the kernels that run color DPs copy it to LDS
// (copy_color_table) and look entries up there.
constexpr int COLOR_TAB = 1024;

constexpr uint8_t color_diag_entry(int idx) {
  const int c_chk = idx & 3, c_auto = (idx >> 2) & 3, c_fresh = (idx >> 4) & 3;
  const int chk0 = (idx >> 6) & 3, q = (idx >> 8) & 3;
  const int x = c_chk ^ chk0;   // COMPOSE(crt checksum, corner checksum)
  int chk = x, autoc = 0, fresh = 0;
  if (chk0 != 0) {              // the corner's checksum changes: a color mismatch in this cell
    if (c_chk == 0) { autoc = x; fresh = q; }                  // a SNP starts or a read error
    else if (x == c_auto) { chk = 0; autoc = 0; fresh = 0; }   // the SNP ends
    else if (x == 0) {
      if (c_fresh - 1 > 0) { autoc = c_auto; fresh = c_fresh - 1; }
      else { autoc = 0; fresh = 0; }
    } else {
      if (c_fresh - 1 > 0) { autoc = c_auto; fresh = c_fresh - 1; }
      else { autoc = x; fresh = q; }
    }
  } else {                      // match: the pending mismatch wears off (else the cell's own
    if (c_auto != 0) {          // initial state: auto = chk0 = 0, fresh = 0)
      if (c_fresh - 1 > 0 && x != 0) { autoc = c_auto; fresh = c_fresh - 1; }
    } else {
      chk = 0;
    }
  }
  const int refund_ok = (c_chk != 0) && (x == 0 || x == c_auto);
  return (uint8_t)(chk | (autoc << 2) | (fresh << 4) | (refund_ok << 6));
}

struct alignas(16) ColorTable {
  uint8_t v[COLOR_TAB];
};
constexpr ColorTable make_color_table() {
  ColorTable t{};
  for (int e = 0; e < COLOR_TAB; ++e) t.v[e] = color_diag_entry(e);
  return t;
}
__device__ constexpr ColorTable kColorTable = make_color_table();

// copies the table to LDS (16-byte aligned) with the nt threads of a workgroup; the caller
// synchronises before the first lookup
__device__ __forceinline__ void copy_color_table(uint8_t* lds, int tid, int nt) {
  for (int e = 16 * tid; e < COLOR_TAB; e += 16 * nt)
    *(uint4*)(lds + e) = *(const uint4*)(kColorTable.v + e);
}

// The DP of one candidate (read strand rs = 2*read + strand, diagonal h) as a state machine over
// rows.  The 2I+1 band columns of the current row, the genome planes and the read's codes and
// quality levels are held in registers: bit c of the row-i genome words is code win + i + c; per
// plane NW words from word index wi, a new word fetched every 32 rows, one word ahead of its use,
// so a row issues no dependent loads.
// Directions (TB): the cell (i, c) is nibble c & 7 of word i * DW + (c >> 3) of the lane; the
// lanes of a wavefront interleave their words (tbw = wavefront base + lane, word k at tbw[64 k]),
// so each row is DW coalesced 256-byte stores of whole cache lines.
template <int IM, bool COLOR>
struct DPLane {
  static constexpr int NCM = 2 * IM + 1;
  static constexpr int NW = NCM > 32 ? 4 : 3;
  static constexpr int DW = (NCM + 7) / 8;
  int sc[NCM];
  int meta[NCM];
  uint32_t gl[NW], gh[NW], gm[NW];
  int64_t wi;
  int sh;
  const uint32_t* pl;
  const uint32_t* qp;
  uint32_t rlo, rhi, rq;
  uint32_t rlo_n, rhi_n, rq_n;   // the next read and quality words, loaded a word ahead
  int has_qual;
  int L, i, win, S, bp;
  const uint8_t* ctab;   // color mode: the diagonal-move table in LDS (COLOR_TAB entries)

  // row 0 (alignment.c:326-360); tbw (TB) receives the row's directions
  template <bool TB>
  __device__ void begin(const GenomeDev& g, const ReadsDev& r, const DPConst& k, int rs, int h, uint32_t* tbw,
                        const uint8_t* color_tab) {
    ctab = color_tab;
    L = r.len[rs >> 1];
    win = (h - k.I > -k.d) ? h - k.I : -k.d;
    pl = read_planes(r, rs);
    qp = r.qual + (int64_t)rs * QW;
    has_qual = r.has_qual;
    const int64_t wbase = (int64_t)win + GPAD;
    wi = wbase >> 5;
    sh = (int)(wbase & 31);
#pragma unroll
    for (int q = 0; q < NW; ++q) {
      const uint2 p = g.lh[wi + q];
      gl[q] = p.x; gh[q] = p.y; gm[q] = g.mask[wi + q];
    }
    rlo = pl[0];
    rhi = pl[RW];
    rq = has_qual ? qp[0] : 0xffffffffu;
    rlo_n = L > 32 ? pl[1] : 0u;
    rhi_n = L > 32 ? pl[RW + 1] : 0u;
    rq_n = has_qual && L > 16 ? qp[1] : 0xffffffffu;
    const int NC = k.NC, I = k.I;
    const int r0 = (int)((rlo & 1u) | ((rhi & 1u) << 1)), q0 = (int)(rq & 3u);
    const int s_m = k.match(q0), s_x = k.mismatch(q0);
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
            meta[c] = m_make(3, 0, chk, chk, chk ? q0 : 0) | (k.m_diag & ~0xFFFF);
          } else {
            sc[c] = (r0 != gj) ? s_x : s_m;
            meta[c] = k.m_diag;
          }
        } else {
          sc[c] = MIN_SCORE;
          meta[c] = k.m_diag;
        }
      } else {   // columns past the band hold the border, so the top neighbour needs no test
        sc[c] = MIN_SCORE;
        meta[c] = m_make(3, I, 0, 0, 0);
      }
    }
    if (TB) {   // row 0: every cell from the corner (3)
#pragma unroll
      for (int w = 0; w < DW; ++w) tbw[64 * w] = 0x33333333u;
    }
    bp = (COLOR ? (L - 2) : L) * k.top_match;
    S = 0x7fffffff;
    i = 0;
  }

  // true while rows remain (row i + 1 < L)
  __device__ bool more() const { return i + 1 < L; }

  // the next row (i + 1) of the recurrence, and the survival bound after it
  template <bool TB>
  __device__ void row(const GenomeDev& g, const DPConst& k, uint32_t* tbw) {
    ++i;
    // advance the genome windows by one code
    if (++sh == 32) {
      sh = 0;
      ++wi;
#pragma unroll
      for (int q = 0; q + 1 < NW; ++q) { gl[q] = gl[q + 1]; gh[q] = gh[q + 1]; gm[q] = gm[q + 1]; }
      const uint2 p = g.lh[wi + NW - 1];
      gl[NW - 1] = p.x;
      gh[NW - 1] = p.y;
      gm[NW - 1] = g.mask[wi + NW - 1];
    }
    // read and quality words: the one loaded a word ahead, and the load of the next (so a row
    // never waits for memory on them)
    if ((i & 31) == 0) {
      rlo = rlo_n;
      rhi = rhi_n;
      if (i + 32 < L) { rlo_n = pl[(i >> 5) + 1]; rhi_n = pl[RW + (i >> 5) + 1]; }
    }
    if ((i & 15) == 0 && has_qual) {
      rq = rq_n;
      if (i + 16 < L) rq_n = qp[(i >> 4) + 1];
    }
    const int NC = k.NC, I = k.I, I4 = k.I4, gap_open = k.gap_open, gap_ext = k.gap_ext;
    const int ri = (int)(((rlo >> (i & 31)) & 1u) | (((rhi >> (i & 31)) & 1u) << 1));
    const int qi = (int)((rq >> ((i & 15) << 1)) & 3u);
    const int s_m = k.match(qi), s_x = k.mismatch(qi);
    // color: a pair score after the correction refund (-2 x a negative score, alignment.c:465-468)
    const int r_m = COLOR ? (s_m < 0 ? -2 * s_m : s_m) : 0, r_x = COLOR ? (s_x < 0 ? -2 * s_x : s_x) : 0;
    const int q12 = qi << 12;
    const uint32_t wl0 = __builtin_amdgcn_alignbit(gl[1], gl[0], sh), wh0 = __builtin_amdgcn_alignbit(gh[1], gh[0], sh);
    const uint32_t wm0 = __builtin_amdgcn_alignbit(gm[1], gm[0], sh);
    const uint32_t wl1 = NCM > 32 ? __builtin_amdgcn_alignbit(gl[2], gl[1], sh) : 0u;
    const uint32_t wh1 = NCM > 32 ? __builtin_amdgcn_alignbit(gh[2], gh[1], sh) : 0u;
    const uint32_t wm1 = NCM > 32 ? __builtin_amdgcn_alignbit(gm[2], gm[1], sh) : 0u;
    // bit c set when read code ri differs from genome code win + i + c
    const uint32_t rb0 = 0u - (uint32_t)(ri & 1), rb1 = 0u - (uint32_t)((ri >> 1) & 1);
    const uint32_t mis0 = (wl0 ^ rb0) | (wh0 ^ rb1), mis1 = NCM > 32 ? ((wl1 ^ rb0) | (wh1 ^ rb1)) : 0u;
    int max_on_line = MIN_SCORE;
    // the left neighbour: its meta word and the score a horizontal gap step out of it gives
    int left_try = MIN_SCORE, left_meta = m_make(3, I, 0, 0, 0);   // border (alignment.c:303-318)
    uint32_t dw[DW];
#pragma unroll
    for (int w = 0; w < DW; ++w) dw[w] = 0;
    // the cells of the row, in one of three copies: a row in which no lane's band touches a masked
    // genome code (N runs: rare) takes a copy without the masked-cell selects, and when the band
    // has the kernel's full width (I = IM, the usual case) one without the per-column width test,
    // so the cells form one block the compiler schedules as a whole
    auto cells = [&](auto mk, auto full) {
      constexpr bool MK = decltype(mk)::value, FULL = decltype(full)::value;
#pragma unroll
      for (int c = 0; c < NCM; ++c) {
        if (FULL || c < NC) {
          const uint32_t Lw = c < 32 ? wl0 : wl1, Hw = c < 32 ? wh0 : wh1, Kw = c < 32 ? wm0 : wm1;
          const int gj = (int)(((Lw >> (c & 31)) & 1u) | (((Hw >> (c & 31)) & 1u) << 1));
          const bool masked = MK && ((Kw >> (c & 31)) & 1u) != 0;
          // this cell before the recurrence (init_matrix): pair score, or MIN_SCORE when masked;
          // in color mode the checksum, auto-correction and freshness of a mismatch
          const bool mis = (((c < 32 ? mis0 : mis1) >> (c & 31)) & 1u) != 0;
          const int cchk = MK ? msel(masked, 0, ri ^ gj) : (ri ^ gj);
          const int cs = MK ? msel(masked, MIN_SCORE, msel(mis, s_x, s_m)) : msel(mis, s_x, s_m);
          const int csc = sc[c], cmeta = meta[c];                          // corner (i-1, j-1)
          const int tsc = (c + 1 < NCM) ? sc[(c + 1) < NCM ? c + 1 : c] : MIN_SCORE;   // top (i-1, j)
          const int tmeta = (c + 1 < NCM) ? meta[(c + 1) < NCM ? c + 1 : c] : m_make(3, I, 0, 0, 0);
          // The recurrence in select form (alignment.c:465-604): every lane runs the same
          // instructions whatever its record, with no divergent branches per cell.
          int match_score = cs;
          // color state of the corner, kept in place in its meta word (chk bits 8-9, auto 10-11,
          // fresh 12-13): the diagonal move's new state and refund from the table
          const int dtab = COLOR ? (int)ctab[((cmeta >> 8) & 0x3F) | (cchk << 6) | (qi << 8)] : 0;
          const int iauto_w = COLOR ? (cmeta & 0xC00) : 0;
          if (COLOR) {   // alignment.c:465-468; & (not &&): no divergent branches
            if (MK) {
              const bool refund = ((dtab & 0x40) != 0) & (cs != MIN_SCORE) & (cs < 0);
              match_score = msel(refund, -2 * cs, cs);
            } else {   // no masked cell: the refunded score of the cell's pair, where the table allows it
              match_score = msel((dtab & 0x40) != 0, msel(mis, r_x, r_m), cs);
            }
          }
          const int diag = csc + match_score;
          const int t_lo8 = tmeta & 0xff, l_lo8 = left_meta & 0xff;
          // a cell's from is 1 (left), 2 (top) or 3 (corner; also row 0 and the band border): a gap
          // extends from the top cell iff its from is 2, from the left iff 1; the top cell carries
          // its vertical step's penalty in its meta word, the left one's step was scored with it
          const int t_try = tsc + (int)(short)(tmeta >> 16);
          const bool take_t = (t_lo8 < I4) & (t_try > diag);   // indels < I

          const int after_t = msel(take_t, t_try, diag);
          const int l_try = left_try;
          const bool take_l = (l_lo8 < I4) & (l_try > after_t);
          const int nsc = msel(take_l, l_try, after_t);
          // low byte and vertical penalty of the new meta word: the corner's indels with from = 3, or
          // the gap neighbour's indels + 1 with from = 2 (top) / 1 (left) (indels <= 31, so +4 never
          // carries out of the byte), and the penalty of a vertical step out of this cell
          const bool gap = take_l | take_t;
          const int nb = msel(take_l, left_meta, tmeta);   // the gap neighbour
          const int d_lo = (cmeta & 0xff) | k.m_diag;
          const int g_lo = (nb & 0xfc) + msel(take_l, k.m_left, k.m_top);
          const int n_lo = msel(gap, g_lo, d_lo);
          const int nfrom = n_lo & 3;
          int nhi = 0;   // color state bits 8-13 of the new meta word
          if (COLOR) {
            const int d_hi = (dtab & 0x3F) << 8;   // diagonal move (alignment.c:533-578)
            // gap from the top (alignment.c:580-589) or from the left (alignment.c:590-599): chk of
            // the neighbour xor the read / genome code; auto of the corner and the neighbour's
            // freshness - 1 while it is > 1, else the cell's own initial state (auto = chk0, fresh =
            // the level when chk0 != 0)
            const bool c_mis = MK ? (cchk != 0) : mis;
            const int c_af = msel(c_mis, (cchk << 10) | q12, 0);
            const int code = msel(take_l, gj, ri);
            const int nfw = nb & 0x7000;
            const int g_hi = ((nb ^ (code << 8)) & 0x300) | msel(nfw > 0x1000, iauto_w | (nfw - 0x1000), c_af);
            nhi = msel(gap, g_hi, d_hi);
          }
          // direction nibble: from, and in color mode whether the score rose (the walk's
          // correction test, alignment.c:684-705)
          if (TB) dw[c >> 3] |= (uint32_t)(nfrom | ((COLOR && nsc > csc) ? 4 : 0)) << ((c & 7) * 4);
          sc[c] = nsc;
          meta[c] = COLOR ? (n_lo | nhi) : n_lo;
          left_try = nsc + msel(take_l, gap_ext, gap_open);   // a horizontal step out of this cell
          left_meta = meta[c];
          if (max_on_line < nsc) max_on_line = nsc;
        }
      }
    };
    if constexpr (NCM > 16) {   // wide bands: one copy (three would spill registers)
      cells(std::true_type{}, std::false_type{});
    } else {
      if (__builtin_amdgcn_ballot_w64((wm0 | wm1) != 0u)) cells(std::true_type{}, std::false_type{});
      else if (NC == NCM) cells(std::false_type{}, std::true_type{});
      else cells(std::false_type{}, std::false_type{});
    }
    if (TB) {
#pragma unroll
      for (int w = 0; w < DW; ++w) tbw[64 * (i * DW + w)] = dw[w];
    }
    if (max_on_line + bp < S) S = max_on_line + bp;
    bp -= k.top_match;
  }

  // score pass result after an early stop (S < t): never accepted
  __device__ AlignRec rejected() const {
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

  // after the last row: the leftmost maximum of the last row (alignment.c:621-630)
  __device__ AlignRec result(const DPConst& k, int& bc) const {
    int best = MIN_SCORE;
    bc = -1;
#pragma unroll
    for (int c = 0; c < NCM; ++c)
      if (c < k.NC && best < sc[c]) { best = sc[c]; bc = c; }
    AlignRec o;
    o.score = best;
    o.S = S;
    o.win = win;
    o.best_j = (int16_t)(bc >= 0 ? L - 1 + bc : -1);
    o.best_i0 = -1;
    o.best_j0 = -1;
    o.tb_len = 0;
    return o;
  }
};

// Nucleotide mode: true when the DP of candidate (rs, h) cannot score above the filter's u8 score
// fscore, so a record with fscore <= t can never be accepted (score > T >= t is required) and the
// DP may be skipped.  Holds when every quality level of the read strand is 3 (the DP's pair scores
// are then the filter's +m / -x; at lower levels its mismatch costs less), I <= 15 and the
// window is not clamped at the sequence start (the DP band, diagonals h-I .. h+I, then lies inside
// the filter's 2W diagonals from h-d, d = I), and fscore < 255 (the filter saturates): the filter
// is a local alignment over a superset of the DP's paths with gap costs never above the DP's
// (one merged gap state, no indel cap) and no masked cells, so its maximum bounds every glocal
// DP path's score.
__device__ __forceinline__ bool dp_bounded_by_filter(const ReadsDev& r, const DPConst& k, int rs, int h) {
  if (k.I > 15 || !(h - k.I > -k.d)) return false;
  if (!r.has_qual) return true;
  const int L = r.len[rs >> 1];
  const uint32_t* qp = r.qual + (int64_t)rs * QW;
  for (int w = 0; w * 16 < L; ++w) {
    const int n = L - w * 16 < 16 ? L - w * 16 : 16;
    const uint32_t m = n == 16 ? 0xffffffffu : ((1u << (2 * n)) - 1u);
    if ((qp[w] & m) != m) return false;
  }
  return true;
}

// The traceback walk of one record (alignment.c:642-734) from (L-1, bc) over its direction words
// (nibble c & 7 of word i * DW + (c >> 3), at tbw[STRIDE * (i * DW + (c >> 3))]: the lanes of a
// wavefront interleave their words), codes to tb in the reference's packing (bytes past the codes'
// last 32-bit word are left as they were; the rest of that word is 0).  Each step reads a
// word the previous step chose, so the rows are loaded RB at a time (one memory latency per RB
// rows instead of one per step) and the step picks its word from registers; the read and genome
// words the steps compare are kept in registers too and loaded again only when the walk leaves
// them (every 32 rows or columns), not at every step.
// LSTRIDE > 0: the block of rows sits in LDS instead of registers (the lane's words at
// lrows[LSTRIDE * (q * DW + w)], lanes interleaved), so a step reads its word with one LDS load
// rather than selecting it among RB x DW registers.
template <int IM, bool COLOR, int STRIDE, int LSTRIDE = 0>
__device__ __forceinline__ void tb_walk(const GenomeDev& g, const uint32_t* pl, int L, int win, int NC, int bc,
                                        const uint32_t* tbw, uint8_t* tb, AlignRec& o, uint32_t* lrows = nullptr) {
  constexpr int DW = (2 * IM + 1 + 7) / 8;
  constexpr int RB = DW <= 2 ? 8 : DW <= 4 ? 4 : 2;
  uint32_t rows[LSTRIDE > 0 ? 1 : RB][DW];
  int lo = L;   // rows lo .. lo + RB - 1 are in `rows`
  int rw = -1;   // read word rw (codes 32 rw ..) in rl / rh
  uint32_t rl = 0, rh = 0;
  int64_t gw = -1;   // genome word gw (codes 32 gw - GPAD ..) in gwp
  uint2 gwp = make_uint2(0u, 0u);
  uint32_t* tbw32 = (uint32_t*)tb;   // tb is 4-byte aligned (records of TBB bytes)
  uint32_t acc = 0;
  int i = L - 1, c = bc, n = 0, color = 0;
  while (i >= 0 && c >= 0 && c < NC) {
    if (i < lo) {   // the RB rows up to row i, loaded together (rows below 0 are not read)
      lo = i - RB + 1;
      if constexpr (LSTRIDE > 0) {
        uint32_t v[RB][DW];
#pragma unroll
        for (int q = 0; q < RB; ++q)
#pragma unroll
          for (int w = 0; w < DW; ++w) v[q][w] = lo + q >= 0 ? tbw[STRIDE * ((lo + q) * DW + w)] : 0u;
#pragma unroll
        for (int q = 0; q < RB; ++q)
#pragma unroll
          for (int w = 0; w < DW; ++w) lrows[LSTRIDE * (q * DW + w)] = v[q][w];
      } else {
#pragma unroll
        for (int q = 0; q < RB; ++q)
#pragma unroll
          for (int w = 0; w < DW; ++w) rows[q][w] = lo + q >= 0 ? tbw[STRIDE * ((lo + q) * DW + w)] : 0u;
      }
    }
    uint32_t word = 0;
    if constexpr (LSTRIDE > 0) {
      word = lrows[LSTRIDE * ((i - lo) * DW + (c >> 3))];
    } else {
#pragma unroll
      for (int q = 0; q < RB; ++q)
#pragma unroll
        for (int w = 0; w < DW; ++w)
          if (q == i - lo && w == (c >> 3)) word = rows[q][w];
    }
    o.best_i0 = (int16_t)i;
    o.best_j0 = (int16_t)(i + c);
    const int cell = (int)(word >> ((c & 7) * 4)) & 0xF;
    const int from = cell & 3;
    if ((i >> 5) != rw) {
      rw = i >> 5;
      rl = pl[rw];
      rh = pl[RW + rw];
    }
    const int ri = (int)(((rl >> (i & 31)) & 1u) | (((rh >> (i & 31)) & 1u) << 1));
    int code;
    if (from == 2) code = 0xC | ri;             // insertion in the read
    else if (from == 1) code = 0x1;             // deletion
    else {
      const int64_t b = (int64_t)win + i + c + GPAD;
      if ((b >> 5) != gw) {
        gw = b >> 5;
        gwp = g.lh[gw];
      }
      const int gj = (int)(((gwp.x >> (b & 31)) & 1u) | (((gwp.y >> (b & 31)) & 1u) << 1));
      if (ri != gj) {
        if (!COLOR) code = 0x4 | ri;
        else if (color || (i != 0 && (cell & 4))) { code = 0x8 | ri; color ^= (ri ^ gj); }
        else code = 0x4 | ri;
      } else code = 0x3;
    }
    // TRACEBACK_SEQ_SET (alignment.h:83): code n in byte n / 2, even index in the high nibble;
    // eight codes make a 32-bit word, stored whole
    acc |= (uint32_t)code << (((n >> 1) & 3) * 8 + ((n & 1) ? 0 : 4));
    if ((n & 7) == 7) {
      tbw32[n >> 3] = acc;
      acc = 0;
    }
    ++n;
    if (from == 3) { --i; }
    else if (from == 2) { --i; ++c; }
    else { --c; }
  }
  if (n & 7) tbw32[n >> 3] = acc;   // the last word, its unused nibbles 0
  o.tb_len = (int16_t)n;
}

// One candidate: read strand rs (2*read + strand), diagonal h.
// TB = true: tbw is this lane's direction words (L rows of DPLane::DW words, interleaved with the
// other lanes of the wavefront) and, when the record can be accepted, the traceback (alignment
// order, 2 codes per byte) goes to tb.
// TB = false: the score pass alone, nothing is stored.
// A record can only ever be accepted when score > T and S >= T with T >= t (T starts at -t and
// only rises, run.c:568,683-685), so both stop at the first row where the bound drops below t
// (S < t then holds) and return score = MIN_SCORE; otherwise they return the same score, S and
// best_j.
template <int IM, bool COLOR, bool TB = true, int LSTRIDE = 0>
__device__ AlignRec dp_one(const GenomeDev& g, const ReadsDev& r, const ScoreDev& scd, int rs, int h, uint32_t* tbw,
                           uint8_t* tb, const uint8_t* color_tab, uint32_t* walk_lds = nullptr) {
  const DPConst k(scd);
  DPLane<IM, COLOR> st;
  st.template begin<TB>(g, r, k, rs, h, tbw, color_tab);
  while (st.more()) {
    st.template row<TB>(g, k, tbw);
    if (st.S < k.t) return st.rejected();   // the rest of the rows cannot raise S
  }
  int bc;
  AlignRec o = st.result(k, bc);
  // A record can be accepted only if score > T and S >= T with T >= t (run.c:568,675-685).
  if (TB && bc >= 0 && o.score > k.t && st.S >= k.t)
    tb_walk<IM, COLOR, 64, LSTRIDE>(g, st.pl, st.L, st.win, k.NC, bc, tbw, tb, o, walk_lds);
  return o;
}

}  // namespace storm
