/*
 * ref_align_harness.c -- drives the reference's own banded DP and traceback (test infrastructure
 * only: the CPU baseline of the DP stage).
 *
 * Linked against alignment.c of the reference where it lies (oracle/Makefile, output in
 * oracle/_ref/libref_align.so, nucleotide build).  For each record it does what process_read does
 * for a filter pass (run.c:665-679): alignment__reset_with_compressed_ref on the window
 * max(h - i, -min(i, 15)), alignment__align with the threshold t, and alignment__traceback when the
 * score exceeds t -- with one AlignmentType per OpenMP thread (run.c:350-357) and the read set with
 * alignment__reset_with_compressed_read (run.c:905).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#include "alignment.h"

/* genome: code 0 of a zero-padded packed sequence; mask: code 0 of its packed masked buffer (1 =
 * masked, padding 0x55); reads: n packed reads of length L, read_stride bytes apart, no quality
 * (every level 3, alignment.c:151-162); record k aligns read k at diagonal diag[k].  Runs the
 * records `reps` times on nthreads threads and returns the wall seconds of one pass; the last pass
 * writes per record the score, the last-row column of the maximum, the traceback start and its
 * codes (tb_stride bytes per record, alignment.h:81-83 packing). */
double ref_align_time(const uint8_t *genome, const uint8_t *mask, const uint8_t *reads, long read_stride, int L,
                      const int *diag, long n, int I, int t, int match, int mismatch, int gap_open, int gap_extend,
                      int nthreads, int reps, int *score, int *best_j, int *best_i0, int *best_j0, int *tb_len,
                      uint8_t *tb, int tb_stride) {
  const int d = I > 15 ? 15 : I;
  double t0 = 0, t1 = 0;
  for (int rep = 0; rep < reps; ++rep) {
    t0 = omp_get_wtime();
#pragma omp parallel num_threads(nthreads)
    {
      AlignmentType al;
      alignment__init(&al, (short)L, (short)(L + 2 * I), (short)I);
      alignment__set_params(&al, (short)match, (short)mismatch, (short)gap_open, (short)gap_extend);
#pragma omp for schedule(dynamic, 256)
      for (long k = 0; k < n; ++k) {
        alignment__reset_with_compressed_read(&al, (CODE_TYPE *)(reads + k * read_stride), NULL);
        const int win = diag[k] - I > -d ? diag[k] - I : -d;
        alignment__reset_with_compressed_ref(&al, (CODE_TYPE *)genome, (CODE_TYPE *)mask, win);
        const int s = alignment__align(&al, (ScoreType)t);
        int len = 0;
        if (s > t) {
          alignment__traceback(&al);
          len = al.traceback_seq_len;
        }
        if (rep == reps - 1) {
          score[k] = s;
          best_j[k] = al.best_j;
          best_i0[k] = s > t ? al.best_i0 : -1;
          best_j0[k] = s > t ? al.best_j0 : -1;
          tb_len[k] = len;
          if (len) memcpy(tb + k * (long)tb_stride, al.traceback, (size_t)((len + 1) / 2 < tb_stride ? (len + 1) / 2 : tb_stride));
        }
      }
      alignment__destroy(&al);
    }
    t1 = omp_get_wtime();
  }
  return t1 - t0;
}
