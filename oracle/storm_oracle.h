/*
 * storm_oracle.h -- CPU restatement of the SToRM mapping hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker: it is called from
 * tests/, from __graft_entry__.smoke() and from bench.py's cpu_baseline leg,
 * and never from the product (storm_amd/).  It restates, in plain C, the
 * reference algorithm of laurentnoe/storm for
 *   seed keys         (seed.c:203-214),
 *   the CSR index     (index.c:19-241),
 *   candidate diagonals (run.c:470-747),
 *   the u8 banded filter (alignment_simd.c:3731-4113, geometry run.c:15-145),
 *   the banded DP + traceback (alignment.c:326-734),
 *   the hit map       (hit_map.c:50-144),
 *   the -M SAM writer (hit_map.c:197-504, alignment.c:749-887),
 * plus the loaders those stages read from (reads_data.c, reference_data.c,
 * read_quality.c, codes.h/.c).  Both nucleotide and color-space modes are
 * runtime flags here (the reference selects them with -DNUCLEOTIDES).
 *
 * Pinning (how each function is checked; see tests/test_oracle_*.py):
 *   - every function is pinned end to end by SAM files the compiled reference
 *     wrote on seeded inputs (tests/golden/), nucleotide and color, -i 3/7,
 *     -M 1/4;
 *   - the filter is pinned per candidate by u8 scores extracted from the
 *     reference SSE2 routine with a threshold sweep (tests/golden/filter_*),
 *     and by closed forms (perfect match = min(255, 5L), all-mismatch = 0,
 *     no-gap reduction);
 *   - seed keys, CSR buckets and candidate sets by brute force;
 *   - the nucleotide DP by exhaustive enumeration of band paths for tiny L.
 */
#ifndef STORM_ORACLE_H
#define STORM_ORACLE_H

#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MIN_SCORE (-16383)   /* alignment.h:47 */
#define ORC_MAP_DETAIL 4         /* util.h:30 */
#define ORC_UNMAPPED (-1)        /* hit_map.h:9 */
#define ORC_SENSE_FWD 1          /* util.h:40 */
#define ORC_SENSE_REV 2          /* util.h:41 */

typedef struct {
  int weight;        /* number of '#' */
  int length;        /* span */
  int *offsets;      /* offsets of the '#' symbols, ascending */
  int *positions;    /* optional explicit read positions (":p1 p2") */
  int positions_count;
} orc_seed_t;

typedef struct { int start, end; } orc_region_t;

typedef struct {
  char *name;
  int size;
  uint8_t first_base;      /* color mode only */
  long size_shift;         /* padding bytes before/after (reference_data.c:75) */
  uint8_t *seq_alloc, *seq;
  uint8_t *mask_alloc, *mask;
  orc_region_t *regions;
  int n_regions;
} orc_ref_t;

typedef struct {
  char *info;
  uint8_t first_base, first_qual;   /* color mode only */
  uint8_t *qual;                    /* packed 2-bit quality levels, or NULL */
  uint8_t *seq;                     /* packed 2-bit codes, zero padded */
  unsigned long tag;
} orc_read_t;

typedef struct {
  char *name;
  int read_len;
  long n;
  orc_read_t *reads;
} orc_reads_t;

typedef struct {
  int index_size;   /* 4^w */
  int map_size;     /* size - span + 1 */
  int *first;       /* index_size + 1 offsets */
  int *map;         /* positions grouped by key, ascending */
} orc_index_t;

typedef struct {
  int score;
  int ref_id, ref_start, ref_end;
  int read_start, read_end;
  int sense;
  uint8_t *traceback;   /* forward order, 2 codes per byte (alignment.h:81-83) */
  int traceback_len;
} orc_hit_t;

typedef struct {
  int color;              /* 0 nucleotide, 1 color space */
  int match, mismatch, gap_open, gap_extend;   /* signed, as main.c holds them */
  int allowed_indels;     /* -i */
  int simd_diags;         /* min(-i, 15) (main.c:302-313) */
  int t;                  /* -t min_accepted_score */
  int z;                  /* -z filter threshold */
  int qual_base;          /* -b */
  int filter_repeats;     /* index.c:9 */
  int stdev_distance;     /* -l */
  int n_bytes;            /* simd_N_BYTE_table (run.c:146) */
  int sense_mask;         /* -F/-R */
  int ignore_lowercase;   /* -j */
  int unordered_M;        /* -M */
  int uniqueness, uniqueness_threshold;   /* -u */
  int print_all;          /* -A */
  int greedy;             /* -G: greedy ordered mapping (genome_map.c:210-232) */
} orc_params_t;

/* defaults of the reference (alignment.h:24-28,51-52, main.c:37-38) */
void orc_params_default(orc_params_t *p, int color);

/* W: lanes per filter alignment for a given min(i,15) (run.c:109-127) */
int orc_filter_lanes(int simd_diags);
/* N_BYTES for a given min(i,15) (run.c:146) */
int orc_n_bytes(int simd_diags);

/* --- seeds (seed.c:25-214) --- */
int orc_seed_parse_list(const char *src, orc_seed_t **seeds);   /* count, or -1 */
int orc_seed_key(const orc_seed_t *seed, const uint8_t *packed, int pos);
void orc_seeds_free(orc_seed_t *seeds, int n);

/* --- packed code access (codes.h:125-130) --- */
static inline int orc_nth(const uint8_t *s, long n) {
  return (s[n >> 2] >> ((n & 3) << 1)) & 3;
}

/* --- reference + reads loaders --- */
int orc_load_reference(const char *path, int readlen, const orc_params_t *p, orc_ref_t **refs);
void orc_refs_free(orc_ref_t *refs, int n);
long orc_load_reads(const char *path, const char *qual_path, const orc_params_t *p, orc_reads_t *db);
void orc_sort_reads(orc_reads_t *db);
void orc_reads_free(orc_reads_t *db);
void orc_read_reverse(const orc_read_t *src, orc_read_t *dst, int len, int color);

/* --- index (index.c:19-241) --- */
int orc_index_build(const orc_ref_t *ref, const orc_seed_t *seed, const orc_params_t *p, orc_index_t *idx);
void orc_index_free(orc_index_t *idx);
int orc_index_next_hit(const orc_index_t *idx, int key, int *hit_pointer);

/* --- test helpers (oracle_io.c): one FASTA sequence and one seed's index, malloc'd copies --- */
int orc_fasta_sequence(const char *path, int color, int ignore_lowercase, int ref_no, uint8_t **packed, int *size,
                       int **starts, int **ends, int *n_regions);
int orc_fasta_index(const char *path, int color, int ignore_lowercase, int ref_no, const char *seed,
                    int stdev_distance, int **first, int *n_first, int **map, int *n_map);
void orc_free(void *p);

/* --- candidates (run.c:555-747): ascending diagonals, returns count --- */
int orc_candidates(const orc_index_t *idx, const orc_seed_t *seeds, int n_seeds,
                   const uint8_t *read, int read_len, int allowed_indels,
                   int *out, int cap);

/* --- filter (alignment_simd.c:3731-4113): u8 max score of one candidate --- */
int orc_filter_score(const uint8_t *genome, long pos, const uint8_t *read, int read_len, int W,
                     int match, int mismatch, int gap_open, int gap_extend);
unsigned orc_filter_threshold(int z);

/* --- DP (alignment.c:326-734) --- */
typedef struct {
  int score;            /* as alignment__align returns it (MIN_SCORE on early exit) */
  int best_i, best_j, best_i0, best_j0;
  uint8_t tb[1024];     /* traceback, reverse order as alignment__traceback stores it */
  int tb_len;
} orc_align_out_t;

/* read codes/levels unpacked, reference/mask codes unpacked (window of L+2I) */
int orc_align(const uint8_t *read, const uint8_t *qlev, int L,
              const uint8_t *ref, const uint8_t *refmask, int I,
              const orc_params_t *p, int min_accepted, orc_align_out_t *out);
/* survival bound S = min over rows >= 1 of (max_on_line + best_possible) (no T) and the
 * unconditional score; used to check the T-free GPU formulation */
void orc_align_bound(const uint8_t *read, const uint8_t *qlev, int L,
                     const uint8_t *ref, const uint8_t *refmask, int I,
                     const orc_params_t *p, int *score, int *S);

/* --- hit map (hit_map.c:50-144) --- */
int orc_hitmap_update(orc_hit_t *row /* [4] */, int ref_id, int ref_pos,
                      const orc_align_out_t *al, int score, int sense);

/* --- whole run: reads against references (run.c:764-1017), -M mode --- */
typedef struct {
  long n_reads;
  orc_hit_t **map;     /* per read: NULL or [4] */
  long mapped;
  /* counters */
  long n_candidates, n_filter_pass, n_dp, n_accepted;
} orc_result_t;

int orc_map(const orc_reads_t *db, const orc_ref_t *refs, int n_refs,
            const orc_seed_t *seeds, int n_seeds, const orc_params_t *p, orc_result_t *res);
void orc_result_free(orc_result_t *res);

int orc_traceback_to_cigar(const uint8_t *tb, int tb_len, const uint8_t *ref_seq, long ref_start,
                           int color, uint8_t ref_start_symbol, char *cigar, char *seq);
/* ordered mode (no -M, the reference's default): the genome map built from the hit map, greedy
 * (-G) or contextual (genome_map.c:210-536), and its coordinate-ordered SAM (genome_map.c:551-800) */
void orc_write_sam_ordered(FILE *out, const orc_result_t *res, const orc_reads_t *db,
                           const orc_ref_t *refs, int n_refs, const orc_params_t *p,
                           int argc, char **argv);
/* -O: FASTQ of the reads without any hit (hit_map.c:514-578) */
void orc_write_unmapped_fastq(FILE *out, const orc_result_t *res, const orc_reads_t *db, const orc_params_t *p);
void orc_write_sam(FILE *out, const orc_result_t *res, const orc_reads_t *db,
                   const orc_ref_t *refs, int n_refs, const orc_params_t *p,
                   int argc, char **argv);

#ifdef __cplusplus
}
#endif
#endif
