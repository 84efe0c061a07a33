// k_align.hip -- banded glocal DP with traceback (stage d) for gfx950.
//
// Semantics: alignment__align + alignment__traceback (alignment.c:326-360, :488-632, :642-734)
// on the window max(h - I, -d) of length L + 2I (run.c:667, :855), nucleotide or color space.
// One thread owns one candidate that passed the filter; the 2I+1 band columns of a row live in
// registers (template bound IM on I).  One pass computes the rows, stores each row's directions in
// scratch words (4 bits per cell, interleaved across the lanes of a wavefront) and stops at the
// first row where the record can no longer be accepted; a record that can (score > t and S >= t)
// is walked back by the same thread.  (Scoring first and aligning only the acceptable records
// again costs more: the direction stores are cheap next to the cells, and one pass does every
// cell once.)
//
// The kernel is threshold-free: alongside the score it returns the survival bound
//   S = min over rows i >= 1 of (max_on_line(i) + best_possible(i)),
// so the caller reproduces the reference's early exit for any running threshold T as
// "accepted iff S >= T and score > T" (alignment.c:613-616).  In nucleotide mode score <= S
// always; in color mode the correction refund (alignment.c:465-468) can break that, which is why
// S is carried.
#include "dev_align.h"

namespace storm {

// The record's window and the look-ahead words of its planes lie inside the genome (a record
// from a caller that the library did not build)
__device__ __forceinline__ bool record_in_bounds(const AlignArgs& a, int64_t rec) {
  const int rs = a.rec_rs[rec];
  if (rs < 0 || rs >= 2 * a.r.n) return false;
  const int64_t h = a.rec_diag[rec];
  const int64_t win = (h - a.sc.I > -a.sc.d) ? h - a.sc.I : -a.sc.d;
  return win + a.r.len[rs >> 1] + 2 * a.sc.I + 96 <= a.g.size + GPAD;
}

__device__ __forceinline__ AlignRec invalid_record() {
  AlignRec x;
  x.score = MIN_SCORE;
  x.S = MIN_SCORE;
  x.win = 0;
  x.best_j = -2;
  x.best_i0 = x.best_j0 = -1;
  x.tb_len = 0;
  return x;
}

template <int IM, bool COLOR>
__global__ __launch_bounds__(128) void k_align(AlignArgs a) {
  __shared__ __align__(16) uint8_t ctab[COLOR ? COLOR_TAB : 16];
  // the traceback walk's block of direction rows, per lane (lanes interleaved): 8 KB at most,
  // which leaves the kernel's occupancy to its registers
  constexpr int DW = (2 * IM + 1 + 7) / 8;
  constexpr int RB = DW <= 2 ? 8 : DW <= 4 ? 4 : 2;
  __shared__ uint32_t walk_lds[RB * DW * 128];
  if (COLOR) {
    copy_color_table(ctab, threadIdx.x, blockDim.x);
    __syncthreads();
  }
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.n) return;
  const int64_t rec = a.base + k;
  if (a.check_bounds && !record_in_bounds(a, rec)) {
    a.out[rec] = invalid_record();
    return;
  }
  a.out[rec] = dp_one<IM, COLOR, true, 128>(a.g, a.r, a.sc, a.rec_rs[rec], a.rec_diag[rec],
                                            (uint32_t*)(a.scratch + (k >> 6) * 64 * a.scratch_stride) + (k & 63),
                                            a.tb + rec * TBB, ctab, walk_lds + threadIdx.x);
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
