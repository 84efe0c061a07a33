// k_align.hip -- banded glocal DP with traceback (stage d) for gfx950.
//
// Semantics: alignment__align + alignment__traceback (alignment.c:326-360, :488-632, :642-734)
// on the window max(h - I, -d) of length L + 2I (run.c:667, :855), nucleotide or color space.
// One thread owns one candidate that passed the filter; the 2I+1 band columns of a row live in
// registers (template bound IM on I), the per-cell direction byte goes to a scratch matrix in
// HBM and the same thread walks it back.
//
// The kernel is threshold-free: alongside the score it returns the survival bound
//   S = min over rows i >= 1 of (max_on_line(i) + best_possible(i)),
// so the caller reproduces the reference's early exit for any running threshold T as
// "accepted iff S >= T and score > T" (alignment.c:613-616).  In nucleotide mode score <= S
// always; in color mode the correction refund (alignment.c:465-468) can break that, which is why
// S is carried.
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

template <int IM, bool COLOR>
__global__ __launch_bounds__(128) void k_align(AlignArgs a) {
  constexpr int NCM = 2 * IM + 1;
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.n) return;
  const int64_t rec = a.base + k;
  const int rs = a.rec_rs[rec];
  const int h = a.rec_diag[rec];
  const int L = a.r.len[rs >> 1];
  const int I = a.sc.I, d = a.sc.d;
  const int NC = 2 * I + 1;
  const int win = (h - I > -d) ? h - I : -d;
  const uint32_t* pl = read_planes(a.r, rs);
  uint8_t* tbm = a.scratch + k * a.scratch_stride;
  const int top_match = a.sc.ps[0][3];

  int sc[NCM];
  int meta[NCM];
  // row 0 (alignment.c:326-360)
  {
    const int r0 = rcode(pl, 0), q0 = rqual(a.r, rs, 0);
#pragma unroll
    for (int c = 0; c < NCM; ++c) {
      if (c < NC) {
        const int j = c;
        const int gj = gcode(a.g, (int64_t)win + j);
        if (gmask(a.g, (int64_t)win + j) == 0) {
          if (COLOR) {
            const int chk = r0 ^ gj;
            sc[c] = a.sc.ps[chk ? 1 : 0][q0];
            meta[c] = m_make(3, 0, chk, chk, chk ? q0 : 0);
          } else {
            sc[c] = a.sc.ps[r0 != gj ? 1 : 0][q0];
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
    const int ri = rcode(pl, i), qi = rqual(a.r, rs, i);
    int max_on_line = MIN_SCORE;
    int left_sc = MIN_SCORE, left_meta = m_make(3, I, 0, 0, 0);   // border (alignment.c:303-318)
#pragma unroll
    for (int c = 0; c < NCM; ++c) {
      if (c < NC) {
        const int j = i + c;
        const int gj = gcode(a.g, (int64_t)win + j);
        const bool masked = gmask(a.g, (int64_t)win + j) != 0;
        // this cell before the recurrence (init_matrix)
        int cs, cchk = 0, cauto = 0, cfresh = 0;
        if (!masked) {
          if (COLOR) {
            cchk = ri ^ gj; cauto = cchk; cfresh = cchk ? qi : 0;
            cs = a.sc.ps[cchk ? 1 : 0][qi];
          } else {
            cs = a.sc.ps[ri != gj ? 1 : 0][qi];
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
          const int tmp = tsc + ((m_from(tmeta) ^ 2) ? a.sc.gap_open : a.sc.gap_ext);
          if (tmp > nsc) { it = 1; nsc = tmp; }
        }
        if (m_ind(left_meta) < I) {
          const int tmp = left_sc + ((m_from(left_meta) ^ 1) ? a.sc.gap_open : a.sc.gap_ext);
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
  if (bc >= 0 && best > a.sc.t && S >= a.sc.t) {
    uint8_t* tb = a.tb + rec * TBB;
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
        const int gj = gcode(a.g, (int64_t)win + j);
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
  a.out[rec] = o;
}

hipError_t launch_align(const AlignArgs& a, hipStream_t st) {
  if (a.n == 0) return hipSuccess;
  const int bs = 128;
  const dim3 grid((unsigned)((a.n + bs - 1) / bs));
  const int I = a.sc.I;
  const bool col = a.sc.color != 0;
#define STORM_ALIGN_CASE(IMV)                                                              \
  if (I <= IMV) {                                                                          \
    if (col) hipLaunchKernelGGL((k_align<IMV, true>), grid, dim3(bs), 0, st, a);           \
    else hipLaunchKernelGGL((k_align<IMV, false>), grid, dim3(bs), 0, st, a);              \
    return hipGetLastError();                                                              \
  }
  STORM_ALIGN_CASE(1)
  STORM_ALIGN_CASE(3)
  STORM_ALIGN_CASE(7)
  STORM_ALIGN_CASE(15)
  STORM_ALIGN_CASE(31)
#undef STORM_ALIGN_CASE
  return hipErrorInvalidValue;
}

}  // namespace storm
