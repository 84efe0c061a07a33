/*
 * ref_filter_harness.c -- drives the reference's own SIMD filter routines (test infrastructure only).
 *
 * Linked against alignment_simd.c of the reference where it lies (oracle/Makefile, output in
 * oracle/_ref/libref_filter.so).  It calls alignment_<isa>__init_<w> / __align_<w> exactly as
 * run.c does (run.c:660, init run.c:839), one read per call with simd_mul window starts, and offers
 *   - pass bits at a threshold,
 *   - exact u8 scores by a threshold sweep (bit(t) = [score > t] is monotone in t, so
 *     score = #{t in [0,254] : bit(t)}),
 *   - an OpenMP timing loop used as the reference CPU baseline of the filter.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#include "alignment_simd.h"

typedef unsigned (*align_fn)(unsigned char *, int *, unsigned char *);
typedef void (*init_fn)(unsigned, unsigned, unsigned, unsigned, unsigned, unsigned);

/* run.c:19/64/109: candidates per call for d = min(i,15) */
static const int mul_sse2[16] = {8, 8, 4, 4, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1};
static const int mul_avx2[16] = {16, 16, 8, 8, 4, 4, 4, 4, 2, 2, 2, 2, 2, 2, 2, 2};
static const int mul_avx512[16] = {32, 32, 16, 16, 8, 8, 8, 8, 4, 4, 4, 4, 4, 4, 4, 4};

static int pick(int isa, int d, align_fn *af, init_fn *inf, void (**clean)(void)) {
  if (isa == 0) {   /* SSE2, the default Makefile build (run.c:105-146) */
    if (!alignment_sse2__compatible_proc()) return -1;
    *clean = alignment_sse2__clean;
    if (d <= 1) { *af = alignment_sse2__align_octa; *inf = alignment_sse2__init_octa; }
    else if (d <= 3) { *af = alignment_sse2__align_quad; *inf = alignment_sse2__init_quad; }
    else if (d <= 7) { *af = alignment_sse2__align_pair; *inf = alignment_sse2__init_pair; }
    else { *af = alignment_sse2__align_mono; *inf = alignment_sse2__init_mono; }
    return mul_sse2[d];
  }
#ifdef __AVX2__
  if (isa == 1) {   /* run.c:60-101 */
    if (!alignment_avx2__compatible_proc()) return -1;
    *clean = alignment_avx2__clean;
    if (d <= 1) { *af = alignment_avx2__align_hexa; *inf = alignment_avx2__init_hexa; }
    else if (d <= 3) { *af = alignment_avx2__align_octa; *inf = alignment_avx2__init_octa; }
    else if (d <= 7) { *af = alignment_avx2__align_quad; *inf = alignment_avx2__init_quad; }
    else { *af = alignment_avx2__align_pair; *inf = alignment_avx2__init_pair; }
    return mul_avx2[d];
  }
#endif
#ifdef __AVX512BW__
  if (isa == 2) {   /* run.c:15-56 */
    if (!alignment_avx512bw__compatible_proc()) return -1;
    *clean = alignment_avx512bw__clean;
    if (d <= 1) { *af = alignment_avx512bw__align_tria; *inf = alignment_avx512bw__init_tria; }
    else if (d <= 3) { *af = alignment_avx512bw__align_hexa; *inf = alignment_avx512bw__init_hexa; }
    else if (d <= 7) { *af = alignment_avx512bw__align_octa; *inf = alignment_avx512bw__init_octa; }
    else { *af = alignment_avx512bw__align_quad; *inf = alignment_avx512bw__init_quad; }
    return mul_avx512[d];
  }
#endif
  return -1;
}

int ref_filter_isa_ok(int isa) {
  align_fn af; init_fn inf; void (*cl)(void);
  return pick(isa, 7, &af, &inf, &cl) > 0;
}

/* Candidates are grouped by read: read r owns candidates [seg[r], seg[r+1]) with window starts
 * pos[]; reads are packed 2-bit codes at reads + r * read_stride.  genome points at code 0 of a
 * zero-padded packed sequence.  Writes one pass bit per candidate. */
static void run_bits(align_fn af, int mul, const uint8_t *genome, long n_reads, const long *seg, const int *pos,
                     const uint8_t *reads, long read_stride, uint8_t *bits, int nthreads) {
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
  for (long r = 0; r < n_reads; ++r) {
    int win[32];
    for (long c0 = seg[r]; c0 < seg[r + 1]; c0 += mul) {
      int n = (int)(seg[r + 1] - c0 < mul ? seg[r + 1] - c0 : mul);
      for (int k = 0; k < mul; ++k) win[k] = pos[c0 + (k < n ? k : n - 1)];   /* run.c:705-711 tail padding */
      unsigned x = af((unsigned char *)genome, win, (unsigned char *)(reads + r * read_stride));
      for (int k = 0; k < n; ++k) bits[c0 + k] = (uint8_t)((x >> k) & 1);
    }
  }
}

int ref_filter_bits(int isa, int d, int L, int match, int mismatch, int gap_open, int gap_extend, int z,
                    const uint8_t *genome, long n_reads, const long *seg, const int *pos,
                    const uint8_t *reads, long read_stride, uint8_t *bits, int nthreads) {
  align_fn af; init_fn inf; void (*cl)(void);
  int mul = pick(isa, d, &af, &inf, &cl);
  if (mul < 0) return -1;
  inf((unsigned)abs(match), (unsigned)abs(mismatch), (unsigned)abs(gap_open), (unsigned)abs(gap_extend), (unsigned)z, (unsigned)L);
  run_bits(af, mul, genome, n_reads, seg, pos, reads, read_stride, bits, nthreads);
  cl();
  return 0;
}

int ref_filter_scores(int isa, int d, int L, int match, int mismatch, int gap_open, int gap_extend,
                      const uint8_t *genome, long n_reads, const long *seg, const int *pos,
                      const uint8_t *reads, long read_stride, uint8_t *scores, int nthreads) {
  align_fn af; init_fn inf; void (*cl)(void);
  int mul = pick(isa, d, &af, &inf, &cl);
  if (mul < 0) return -1;
  long n = seg[n_reads];
  uint8_t *bits = malloc((size_t)(n ? n : 1));
  memset(scores, 0, (size_t)n);
  for (int t = 0; t <= 254; ++t) {
    inf((unsigned)abs(match), (unsigned)abs(mismatch), (unsigned)abs(gap_open), (unsigned)abs(gap_extend), (unsigned)t, (unsigned)L);
    run_bits(af, mul, genome, n_reads, seg, pos, reads, read_stride, bits, nthreads);
    for (long c = 0; c < n; ++c) scores[c] = (uint8_t)(scores[c] + bits[c]);
  }
  cl();
  free(bits);
  return 0;
}

/* Timing loop: `reps` passes over the candidates; returns seconds. */
double ref_filter_time(int isa, int d, int L, int z, const uint8_t *genome, long n_reads, const long *seg,
                       const int *pos, const uint8_t *reads, long read_stride, uint8_t *bits, int nthreads, int reps) {
  align_fn af; init_fn inf; void (*cl)(void);
  int mul = pick(isa, d, &af, &inf, &cl);
  if (mul < 0) return -1.0;
  inf(5, 4, 9, 4, (unsigned)z, (unsigned)L);
  double t0 = omp_get_wtime();
  for (int k = 0; k < reps; ++k) run_bits(af, mul, genome, n_reads, seg, pos, reads, read_stride, bits, nthreads);
  double t1 = omp_get_wtime();
  cl();
  return t1 - t0;
}
