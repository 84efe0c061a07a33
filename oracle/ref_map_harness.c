/*
 * ref_map_harness.c -- times the reference's own mapping phase (test infrastructure only: the CPU
 * baseline of the mapping workloads).
 *
 * Linked with every reference source file where it lies (oracle/Makefile, output in oracle/_ref/),
 * the reference's main renamed so that one process can run it several times:
 *
 *   ref-map-harness-{nucleotide,color} TIMES COMMON... -- RUN1... [-- RUN2... ...]
 *
 * runs the reference's command line COMMON + RUNk once per run (each run reads one read file of a
 * single length, as the reference requires, reads_data.c:156-163) and appends to the file TIMES one
 * line per run: "run <k> map_seconds <t>", t = wall time from the end of the seed index build to
 * the start of the SAM output: the forward and reverse passes of process_read over every read
 * (run.c:865-994) -- what the GPU path replaces -- without the loading, indexing and SAM writing
 * around them.
 *
 * Through the linker's symbol wrapping (-Wl,--wrap) the reference sequences and the seed indexes
 * are built once, by the first run, and handed to the later runs (which the reference would
 * rebuild identically from the same files; the first run must hold the longest reads, since the
 * sequence padding follows the read length, reference_data.c:75-80).
 */
#undef main   /* the reference's main.c is built with main renamed to storm_reference_main */
#define _GNU_SOURCE
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include "hit_map.h"
#include "index.h"
#include "reference_data.h"
#include "seed.h"

int storm_reference_main(int argc, char **argv);

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* ------------------------------------------------------------------ sequences, loaded once */
int __real_load_reference_db(const char *reference_filename, int readlen, ReferenceDBType **db);
void __real_clear_reference_db(ReferenceDBType *db);

static ReferenceDBType *g_refs = NULL;
static int g_n_refs = 0;

int __wrap_load_reference_db(const char *reference_filename, int readlen, ReferenceDBType **db) {
  if (!g_refs) {
    g_n_refs = __real_load_reference_db(reference_filename, readlen, &g_refs);
    if (g_n_refs <= 0) return g_n_refs;
  }
  /* a run frees the array it gets (run.c:1102): it gets a copy of the descriptors */
  *db = (ReferenceDBType *)malloc(sizeof(ReferenceDBType) * (size_t)g_n_refs);
  memcpy(*db, g_refs, sizeof(ReferenceDBType) * (size_t)g_n_refs);
  return g_n_refs;
}

void __wrap_clear_reference_db(ReferenceDBType *db) { (void)db; /* the sequences outlive the run */ }

/* ------------------------------------------------------------------ indexes, built once */
IndexType *__real_index__build_reference(ReferenceDBType *db, SeedType *seed);
void __real_index__destroy(IndexType *index);

#define MAX_CACHED 64
static struct {
  const void *seq;
  unsigned long long mask;
  int length, weight, n_pos;
  int pos[64];
  IndexType *ix;
} g_ix[MAX_CACHED];
static int g_n_ix = 0;
static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;
static double g_index_done = 0, g_sam_start = 0;

static int same_seed(int k, const void *seq, const SeedType *s) {
  if (g_ix[k].seq != seq || g_ix[k].mask != s->mask || g_ix[k].length != s->length || g_ix[k].weight != s->weight)
    return 0;
  const int np = s->positions ? s->positions_count : 0;
  if (g_ix[k].n_pos != np) return 0;
  for (int i = 0; i < np && i < 64; ++i)
    if (g_ix[k].pos[i] != s->positions[i]) return 0;
  return 1;
}

IndexType *__wrap_index__build_reference(ReferenceDBType *db, SeedType *seed) {
  IndexType *ix = NULL;
  pthread_mutex_lock(&g_mu);
  for (int k = 0; k < g_n_ix; ++k)
    if (same_seed(k, db->sequence, seed)) { ix = g_ix[k].ix; break; }
  pthread_mutex_unlock(&g_mu);
  if (!ix) {
    ix = __real_index__build_reference(db, seed);
    pthread_mutex_lock(&g_mu);
    if (g_n_ix < MAX_CACHED) {
      const int k = g_n_ix++;
      g_ix[k].seq = db->sequence;
      g_ix[k].mask = seed->mask;
      g_ix[k].length = seed->length;
      g_ix[k].weight = seed->weight;
      g_ix[k].n_pos = seed->positions ? seed->positions_count : 0;
      for (int i = 0; i < g_ix[k].n_pos && i < 64; ++i) g_ix[k].pos[i] = seed->positions[i];
      g_ix[k].ix = ix;
    }
    pthread_mutex_unlock(&g_mu);
  }
  /* a run frees the descriptor it gets (run.c:989-991): it gets a copy naming its own sequence
   * descriptor and seed */
  IndexType *copy = (IndexType *)malloc(sizeof(IndexType));
  *copy = *ix;
  copy->db = db;
  copy->seed = seed;
  pthread_mutex_lock(&g_mu);
  const double t = now_s();
  if (t > g_index_done) g_index_done = t;
  pthread_mutex_unlock(&g_mu);
  return copy;
}

void __wrap_index__destroy(IndexType *index) { (void)index; /* the cached arrays outlive the run */ }

/* ------------------------------------------------------------------ end of the mapping phase */
void __real_hit_map__generate_SAM_output(const HitMapType *map, const ReadsDBType *reads_db,
                                         const ReferenceDBType *ref_dbs, const int ref_dbs_size, FILE *sam_output,
                                         const int score_rank_max);

void __wrap_hit_map__generate_SAM_output(const HitMapType *map, const ReadsDBType *reads_db,
                                         const ReferenceDBType *ref_dbs, const int ref_dbs_size, FILE *sam_output,
                                         const int score_rank_max) {
  g_sam_start = now_s();
  __real_hit_map__generate_SAM_output(map, reads_db, ref_dbs, ref_dbs_size, sam_output, score_rank_max);
}

int main(int argc, char **argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s TIMES COMMON... -- RUN1... [-- RUN2...]\n", argv[0]);
    return 2;
  }
  FILE *times = fopen(argv[1], "w");
  if (!times) return 2;
  int c0 = 2, c1 = c0;
  while (c1 < argc && strcmp(argv[c1], "--") != 0) ++c1;
  int run = 0;
  for (int r0 = c1 + 1; r0 < argc;) {
    int r1 = r0;
    while (r1 < argc && strcmp(argv[r1], "--") != 0) ++r1;
    char **av = (char **)calloc((size_t)(1 + (c1 - c0) + (r1 - r0) + 1), sizeof(char *));
    int ac = 0;
    av[ac++] = argv[0];
    for (int i = c0; i < c1; ++i) av[ac++] = argv[i];
    for (int i = r0; i < r1; ++i) av[ac++] = argv[i];
    optind = 0;   /* glibc: rescan a new argument vector */
    g_index_done = g_sam_start = 0;
    const int rc = storm_reference_main(ac, av);
    if (rc != 0 || g_sam_start == 0) {
      fprintf(stderr, "run %d failed (rc %d)\n", run, rc);
      return 1;
    }
    fprintf(times, "run %d map_seconds %.6f\n", run, g_sam_start - g_index_done);
    fflush(times);
    free(av);
    ++run;
    r0 = r1 + 1;
  }
  fclose(times);
  return 0;
}
