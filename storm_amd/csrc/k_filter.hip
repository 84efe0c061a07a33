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
// needed.  Per pair of cells the chain is 9-10 packed VALU ops (v_pk_mad_u16, v_pk_add_u16,
// v_pk_max_u16/i16, v_pk_sub_u16 clamp).  Three rewrites keep it short and exact:
//   * the gap state is carried as G = max(subs(I, ext), subs(M, open)) per cell, so a cell's
//     I is max(G_top, G_left) -- one op instead of four;
//   * when open - ext == match and open >= ext (the default scores), G = subs(max(I, X), ext)
//     with X = corner - (match + mismatch) * [mismatch], the cell's score being X + match: one
//     op fewer (dev_filter.h, filter_row FAST);
//   * the upper u8 clamp is dropped: scores are kept unclamped in 16 bits and the final
//     maximum is clamped to 255.  The first cell that would saturate makes the clamped maximum
//     255 and every cell before it is unchanged, so min(255, max) is the reference's result.
// Read and genome codes arrive as bit planes; the mismatch bits of 32 rows of one diagonal come
// from one funnel shift of each genome plane and an xor with the read's planes, packed 16 rows of
// both candidates per register, so a cell takes its bit with a shift and a mask.
#include "dev_filter.h"
#include "storm_kernels.h"

namespace storm {

template <int W>
__global__ __launch_bounds__(256) void k_filter(FilterArgs a) {
  const int64_t pair = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t c0 = 2 * pair;
  if (c0 >= a.n) return;
  const int64_t c1 = (c0 + 1 < a.n) ? c0 + 1 : c0;
  const int ra = a.cand_read[c0], rb = a.cand_read[c1];
  const int La = a.rlen[ra >> a.len_shift], Lb = a.rlen[rb >> a.len_shift];
  const uint32_t* pa = a.rplanes + (int64_t)ra * (2 * RW);
  const uint32_t* pb = a.rplanes + (int64_t)rb * (2 * RW);
  const int64_t xa = (int64_t)a.cand_pos[c0] + GPAD, xb = (int64_t)a.cand_pos[c1] + GPAD;
  const FilterConst k = filter_const(a.mt, a.mm, a.op, a.ex, a.thr);
  int sa, sb;
  filter_pair<W>(a.g, pa, pb, La, Lb, xa, xb, k, sa, sb);
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
