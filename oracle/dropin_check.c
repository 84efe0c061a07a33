/*
 * dropin_check.c -- the reference's own index build and DP against the gfx950 entry points that
 * keep their signatures (include/storm_ref_dropin.h); test infrastructure only.
 *
 * Linked with the reference's sources where they lie (its main renamed) and libstorm_hip.so
 * (oracle/Makefile, output oracle/_ref/dropin-check-{nucleotide,color}).
 *
 *   dropin-check-X GENOME READS QUAL|- SEEDS I T L
 *
 * loads GENOME with load_reference_db and READS (QUAL) with load_reads_db, then for every sequence
 * and seed compares index__build_reference with index_gfx950__build_reference (sizes, first[],
 * every key's positions through index__get_extern_next_hit on both); and for the reads' seed-hit
 * diagonals (and shifted ones) compares alignment__align + alignment__traceback with
 * alignment_gfx950__align_batch, under thresholds T, T + 25 and T - 40 in turn (scores, best cells,
 * traceback codes).  L sets ACCEPTED_STDEV_DISTANCE (<= 0: no repeat filter).  Prints a summary
 * and exits 0 when everything agrees.
 */
#undef main   /* the reference's main.c is built with main renamed to storm_reference_main */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "alignment.h"
#include "index.h"
#include "reads_data.h"
#include "reference_data.h"
#include "seed.h"
#include "storm_ref_dropin.h"

extern int allowed_indels;
extern int simd_allowed_diags;
extern int N_BYTES;
extern int FILTER_REPEATS;
extern int ACCEPTED_STDEV_DISTANCE;
const char *storm_last_error(void);

static int fails = 0;
#define CHECK(cond, ...) do { if (!(cond)) { if (fails++ < 10) { fprintf(stderr, __VA_ARGS__); fprintf(stderr, "\n"); } } } while (0)

int main(int argc, char **argv) {
  if (argc != 8) {
    fprintf(stderr, "usage: %s GENOME READS QUAL|- SEEDS I T L\n", argv[0]);
    return 2;
  }
  const int I = atoi(argv[5]), T = atoi(argv[6]), L_dist = atoi(argv[7]);
  allowed_indels = I;
  simd_allowed_diags = I > 15 ? 15 : I;
  N_BYTES = 16;
  FILTER_REPEATS = L_dist > 0;
  ACCEPTED_STDEV_DISTANCE = L_dist > 0 ? L_dist : 0;
  ReadsDBType reads;
  if (load_reads_db(argv[2], strcmp(argv[3], "-") ? argv[3] : NULL, &reads) <= 0) return 2;
  const int L = reads.read_len;
  ReferenceDBType *refs;
  const int n_refs = load_reference_db(argv[1], L, &refs);
  if (n_refs <= 0) return 2;
  SeedType **seeds;
  const int n_seeds = seed__parse_list(argv[4], &seeds);
  if (n_seeds <= 0) return 2;
  long keys_checked = 0, hits_checked = 0, recs = 0, accepted = 0;
  for (int r = 0; r < n_refs; ++r) {
    IndexType **ix = calloc((size_t)n_seeds, sizeof(IndexType *));
    for (int s = 0; s < n_seeds; ++s) {
      IndexType *a = index__build_reference(&refs[r], seeds[s]);
      IndexType *b = index_gfx950__build_reference(&refs[r], seeds[s]);
      CHECK(b != NULL, "ref %d seed %d: %s", r, s, storm_last_error());
      if (!b) return 1;
      CHECK(a->index_size == b->index_size && a->map_size == b->map_size, "ref %d seed %d: sizes %d/%d vs %d/%d", r, s,
            a->index_size, a->map_size, b->index_size, b->map_size);
      if (a->map_size > 0) {
        for (int k = 0; k <= a->index_size; ++k) {
          CHECK(a->first[k] == b->first[k], "ref %d seed %d: first[%d] %d vs %d", r, s, k, a->first[k], b->first[k]);
          if (k == a->index_size || a->first[k] == a->first[k + 1]) continue;
          int pa = -1, pb = -1, ha, hb;
          do {   /* the key's positions, through the reference's cursor on both indexes */
            ha = index__get_extern_next_hit(a, k, &pa);
            hb = index__get_extern_next_hit(b, k, &pb);
            CHECK(ha == hb, "ref %d seed %d key %d: hit %d vs %d", r, s, k, ha, hb);
            hits_checked++;
          } while (ha >= 0 && hb >= 0);
          keys_checked++;
        }
      }
      index__destroy(a);
      free(a);
      ix[s] = b;
    }
    /* DP: per read (forward), the diagonals of its first seed hits and two shifted copies */
    int cap = 8 * (int)reads.size + 16, n = 0;
    AlignmentType *al_ref = calloc((size_t)cap, sizeof(AlignmentType)), *al_gpu = calloc((size_t)cap, sizeof(AlignmentType));
    for (long q = 0; q < reads.size && n + 8 <= cap; ++q) {
      int diags[8], nd = 0;
      for (int p = 0; p + seeds[0]->length <= L && nd < 2; p += 7) {
        const int key = seed__apply_to_compressed(seeds[0], reads.reads[q].sequence, p);
        int ptr = -1, h;
        while (nd < 2 && (h = index__get_extern_next_hit(ix[0], key, &ptr)) >= 0) diags[nd++] = h - p;
      }
      const int base = nd;
      for (int e = 0; e < base; ++e) { diags[nd++] = diags[e] + 3; diags[nd++] = diags[e] - I - 2; }
      for (int e = 0; e < nd; ++e) {
        const int h = diags[e];
        const int d = I > 15 ? 15 : I;
        const int win = h - I > -d ? h - I : -d;
        if (win + L + 2 * I > refs[r].size) continue;
        for (int side = 0; side < 2; ++side) {
          AlignmentType *x = side ? &al_gpu[n] : &al_ref[n];
          alignment__init(x, (short)L, (short)(L + 2 * I), (short)I);
          alignment__set_params(x, 5, -4, -9, -4);
          alignment__reset_with_compressed_read(x, reads.reads[q].sequence, reads.reads[q].quality);
          alignment__reset_with_compressed_ref(x, refs[r].sequence, refs[r].sequence_masked, win);
        }
        n++;
      }
    }
    const int thr[3] = {T, T + 25, T - 40};
    ScoreType *mins = malloc(sizeof(ScoreType) * (size_t)(n + 1)), *sg = malloc(sizeof(ScoreType) * (size_t)(n + 1));
    AlignmentType **ptrs = malloc(sizeof(AlignmentType *) * (size_t)(n + 1));
    for (int pass = 0; pass < 3; ++pass) {
      for (int k = 0; k < n; ++k) { mins[k] = thr[(pass + k) % 3]; ptrs[k] = &al_gpu[k]; }
      const int rc = alignment_gfx950__align_batch(ptrs, n, mins, 1, sg);
      CHECK(rc == 0, "align batch: %s", storm_last_error());
      for (int k = 0; k < n; ++k) {
        AlignmentType *a = &al_ref[k], *b = &al_gpu[k];
        const int s = alignment__align(a, mins[k]);
        CHECK(s == sg[k], "ref %d record %d threshold %d: score %d vs %d", r, k, mins[k], s, sg[k]);
        if (s != MIN_SCORE) {
          CHECK(a->best_i == b->best_i && a->best_j == b->best_j, "record %d: best cell", k);
          CHECK(alignment__get_score(b, b->best_i, b->best_j) == s, "record %d: best cell score in the matrix", k);
        }
        if (s > mins[k]) {
          alignment__traceback(a);
          accepted++;
          CHECK(a->best_i0 == b->best_i0 && a->best_j0 == b->best_j0 && a->traceback_seq_len == b->traceback_seq_len,
                "record %d: traceback start or length", k);
          for (int c = 0; c < a->traceback_seq_len && c < b->traceback_seq_len; ++c)
            CHECK(TRACEBACK_SEQ_GET(a->traceback, c) == TRACEBACK_SEQ_GET(b->traceback, c), "record %d: code %d", k, c);
        }
        recs++;
      }
    }
    for (int k = 0; k < n; ++k) { alignment__destroy(&al_ref[k]); alignment__destroy(&al_gpu[k]); }
    free(al_ref); free(al_gpu); free(mins); free(sg); free(ptrs);
    for (int s = 0; s < n_seeds; ++s) { index_gfx950__destroy(ix[s]); free(ix[s]); }
    free(ix);
  }
  printf("keys %ld hits %ld records %ld accepted %ld mismatches %d\n", keys_checked, hits_checked, recs, accepted, fails);
  return fails ? 1 : 0;
}
