/*
 * storm_gpu.h -- C ABI of the MI355X (gfx950) implementation of the SToRM mapping hot path.
 *
 * Library: storm_amd/_build/libstorm_hip.so.  Every function returns an int status
 * (STORM_OK = 0, negative on error) and never exits the process; storm_last_error() gives the
 * message of the calling thread's last failure.  Host buffers belong to the caller; device
 * buffers belong to the context or plan that created them.  A context is used by one host
 * thread at a time.
 *
 * Two groups of entry points:
 *   1. drop-in rows for the reference's call sites (same names/signatures pattern as the
 *      reference, gfx950 in place of the ISA), one GPU launch per call;
 *   2. batch entry points that replace the reference's per-read loop (process_read and the
 *      passes around it, run.c:555-1017) for many reads at once.
 */
#ifndef STORM_GPU_H
#define STORM_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define STORM_OK 0
#define STORM_ERR_INVALID (-1)     /* bad argument (mirrors RETURN_INPUT_ERR, load_data.h:8) */
#define STORM_ERR_MEMORY (-2)      /* allocation failed (mirrors RETURN_MEM_ERR, load_data.h:9) */
#define STORM_ERR_DEVICE (-3)      /* HIP runtime error */
#define STORM_ERR_IO (-4)          /* file could not be read or written */

const char *storm_last_error(void);

/* ------------------------------------------------------------------------------------------
 * Mapping parameters: the reference's command-line options (main.c:188-433) that the -M
 * mapping path reads.
 */
typedef struct storm_params {
  int color;              /* 0: nucleotide (storm-nucleotide), 1: color space (storm-color) */
  int match;              /* -m, default +5 (alignment.h:24) */
  int mismatch;           /* -x, default -4 */
  int gap_open;           /* -d, default -9 */
  int gap_extend;         /* -e, default -4 */
  int allowed_indels;     /* -i, default 3; the filter uses min(i, 15) (main.c:302-313) */
  int t;                  /* -t, alignment threshold, default 115 */
  int z;                  /* -z, filter threshold, default 100 */
  int qual_base;          /* -b, default 33 */
  int stdev_distance;     /* -l, repeat filter, default 100; <= 0 disables it */
  int sense_mask;         /* 1 forward (-F), 2 reverse (-R), 3 both (default) */
  int max_hits;           /* -M, 1..4; 0 = the ordered mode (no -M, the reference's default):
                           * reads placed on a genome map, coordinate-sorted SAM (genome_map.c) */
  int uniqueness;         /* -u threshold, 0 = off */
  int print_all;          /* -A */
  int ignore_lowercase;   /* -j */
  int strict_length;      /* 1: reject mixed read lengths like the reference (reads_data.c:156-163) */
  const char *seeds;      /* -s, ';'-separated spaced seeds; NULL = reference defaults (seed.h:29-32) */
  int greedy;             /* -G: ordered mode places each read at its best hit (genome_map.c:186-232)
                           * instead of the contextual placement (genome_map.c:432-536) */
  const char *unmapped_path;   /* -O: FASTQ of the reads without any hit, in read order
                                * (hit_map__generate_unmapped_FASTQ_output, hit_map.c:514-578,
                                * run.c:1078-1086); NULL = none */
} storm_params;

void storm_params_default(storm_params *p, int color);

/* ------------------------------------------------------------------------------------------
 * Group 1: gfx950 row of the reference's SIMD filter table.
 * Replaces alignment_<isa>__init_<w> / __align_<w> / __clean / __compatible_proc
 * (alignment_simd.h:32-300), reached through simd_init_fct_table / simd_fct_table /
 * simd_clean_fct (run.c:15-221).  W = 2/4/8/16 lanes for octa/quad/pair/mono, filtering
 * simd_mul = 8/4/2/1 windows per call exactly as the SSE2 row does (run.c:109).
 * `genome` must point inside a buffer registered with alignment_gfx950__set_genome (the
 * reference passes ref_seq, run.c:660), `read` is a packed read (codes.h:125-130).
 * Threading, as the reference's rows: set_genome and init_<w> run once before the mapping loops
 * and must not race with align calls; align_<w> calls are reentrant and may run from any number of
 * host threads at once (run.c:900 calls them inside an OpenMP loop): each call uses a staging slot
 * and stream of its own.  An align call that fails (no genome registered, wrong row, device error)
 * returns 0; alignment_gfx950__status() then gives the calling thread's error code (STORM_OK after
 * a call that succeeded) and storm_last_error() the message.
 */
int alignment_gfx950__set_genome(const unsigned char *sequence_alloc, size_t alloc_bytes, size_t code0_offset_bytes);
void alignment_gfx950__init_octa(unsigned match, unsigned mismatch, unsigned gapopen, unsigned gapextends,
                                 unsigned threshold, unsigned readlength);
void alignment_gfx950__init_quad(unsigned match, unsigned mismatch, unsigned gapopen, unsigned gapextends,
                                 unsigned threshold, unsigned readlength);
void alignment_gfx950__init_pair(unsigned match, unsigned mismatch, unsigned gapopen, unsigned gapextends,
                                 unsigned threshold, unsigned readlength);
void alignment_gfx950__init_mono(unsigned match, unsigned mismatch, unsigned gapopen, unsigned gapextends,
                                 unsigned threshold, unsigned readlength);
unsigned alignment_gfx950__align_octa(unsigned char *genome, int *pos_genome, unsigned char *read);
unsigned alignment_gfx950__align_quad(unsigned char *genome, int *pos_genome, unsigned char *read);
unsigned alignment_gfx950__align_pair(unsigned char *genome, int *pos_genome, unsigned char *read);
unsigned alignment_gfx950__align_mono(unsigned char *genome, int *pos_genome, unsigned char *read);
void alignment_gfx950__clean(void);
int alignment_gfx950__compatible_proc(void);
/* No counterpart in the reference (its rows cannot fail): the status of the calling thread's last
 * align_<w> call. */
int alignment_gfx950__status(void);

/* ------------------------------------------------------------------------------------------
 * Group 2: batch filter (stage c) over many candidates, with inputs resident in HBM.
 * Replaces the simd_fct_table calls of process_read (run.c:660, :714) for a whole batch.
 *   genome_packed : packed codes of one sequence (code n at bits 2*(n&3) of byte n>>2), size codes
 *   reads_packed  : n_reads packed reads, read_stride bytes apart; read_len[n_reads]
 *   cand_read/cand_pos : per candidate, read index and window start (hit - min(i,15), run.c:651)
 *   W             : 2, 4, 8 or 16 (run.c:109-127 for -i)
 * storm_filter_run launches the kernel `reps` times on the resident data and reports the mean
 * kernel time in ms (HIP events on the plan's stream); storm_filter_fetch copies the u8 scores
 * (0..255, the reference's running maximum) and pass flags (score > min(z,254)) back.
 */
typedef struct storm_filter_plan storm_filter_plan;
int storm_filter_plan_create(int device, const uint8_t *genome_packed, int64_t genome_codes,
                             const uint8_t *reads_packed, int64_t read_stride, const int32_t *read_len, int64_t n_reads,
                             const int32_t *cand_read, const int32_t *cand_pos, int64_t n_cand,
                             int W, int match, int mismatch, int gap_open, int gap_extend, int z,
                             storm_filter_plan **out);
int storm_filter_run(storm_filter_plan *plan, int reps, double *ms_per_launch);
int storm_filter_fetch(storm_filter_plan *plan, uint8_t *scores, uint8_t *pass);
void storm_filter_plan_destroy(storm_filter_plan *plan);

/* ------------------------------------------------------------------------------------------
 * Batch DP + traceback (stage d) over (read strand, diagonal) records, inputs resident in HBM.
 * Replaces alignment__align + alignment__traceback (alignment.c:488-734, declared
 * alignment.h:280-289) as process_read calls them for each filter pass (run.c:665-690,
 * :720-735), for a whole batch.
 *   genome_packed : one sequence, 2-bit codes (code n at bits 2*(n&3) of byte n>>2), genome_codes
 *   genome_mask   : genome_codes bytes, nonzero = masked (ReferenceDBType.mask), or NULL
 *   reads_packed  : n_reads forward-strand reads, read_stride bytes apart, lengths read_len (1..256)
 *   qual          : 2-bit quality levels (0..3) per read packed like the codes, qual_stride bytes
 *                   apart, or NULL (every level 3, as for reads without quality)
 *   first_qual    : color space: per read, the level of the primer-side color, or NULL (0)
 *   rec_rs        : per record, 2 * read + strand; strand 1 is the reverse complement (color:
 *                   the reversed colors), codes.c:429-439, read_quality.c:47-64
 *   rec_diag      : per record, the diagonal h (read code 0 against genome code h); the DP
 *                   window starts at win = max(h - i, -min(i, 15)) (run.c:667)
 *   params        : color, match, mismatch, gap_open, gap_extend, allowed_indels (-i), t
 * storm_align_run runs the batch `reps` times on the resident data (mean time per run in ms,
 * HIP events); storm_align_fetch copies per record a storm_align_result and STORM_TB_BYTES
 * traceback bytes:
 *   score  : alignment__align's result with min_accepted = t: MIN_SCORE (-16383) when some row's
 *            max + best possible score drops below t (alignment.c:613-616), else the leftmost
 *            maximum of the last row
 *   bound  : the minimum over rows >= 1 of that row bound (up to the row that stopped the DP);
 *            for a running threshold T >= t, alignment__align(T) returns score when bound >= T
 *            and MIN_SCORE otherwise
 *   win, best_j : window start; the last row's column j of the maximum (best_i = L - 1), -1 if none
 *   best_i0, best_j0, tb_len, tb : when score > t, where the traceback ends (row 0) and the
 *            traceback in alignment__traceback's order and packing (TRACEBACK_SEQ_SET,
 *            alignment.h:81-83, end of the alignment first); else -1, -1, 0
 */
#define STORM_TB_BYTES 144
typedef struct storm_align_result {
  int32_t score, bound, win;
  int16_t best_j, best_i0, best_j0, tb_len;
} storm_align_result;
typedef struct storm_align_plan storm_align_plan;
int storm_align_plan_create(int device, const uint8_t *genome_packed, const uint8_t *genome_mask, int64_t genome_codes,
                            const uint8_t *reads_packed, int64_t read_stride, const uint8_t *qual, int64_t qual_stride,
                            const uint8_t *first_qual, const int32_t *read_len, int64_t n_reads, const int32_t *rec_rs,
                            const int32_t *rec_diag, int64_t n_rec, const storm_params *params,
                            storm_align_plan **out);
int storm_align_run(storm_align_plan *plan, int reps, double *ms_per_run);
int storm_align_fetch(storm_align_plan *plan, storm_align_result *results, uint8_t *tb);
void storm_align_plan_destroy(storm_align_plan *plan);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the DP + traceback of
 * n_rec records the caller holds in device memory, d_rec_rs / d_rec_diag with the meaning of
 * storm_align_plan_create's rec_rs / rec_diag, against the plan's resident genome and reads (the
 * plan's own records are not used), written to the caller's device buffers d_results (n_rec
 * storm_align_result) and d_tb (n_rec x STORM_TB_BYTES, 4-byte aligned: tracebacks are written in
 * 32-bit words; an unaligned d_tb or d_results is rejected).  The call allocates nothing: the direction
 * words go to a static 16 MiB device array (held by each GPU the library runs on until the process
 * ends), in launches of as many records as it holds (a wavefront of 64 records needs 64 x longest
 * read x ceil((2 dp_band + 1) / 8) x 4 bytes; reads up to 256 codes at any -i fit; more is rejected
 * with STORM_ERR_INVALID).  The array is shared by every call on a GPU: calls that overlap on one
 * GPU race (use one stream, or wait between calls).  A record whose read strand is outside the
 * plan's reads or whose window runs past the padded genome is not aligned: its result is
 * score = bound = MIN_SCORE, best_j = -2, tb_len = 0.
 * stream: a hipStream_t to enqueue on (the call returns at once), or NULL for the plan's stream
 * (the call returns when the results are written). */
int storm_align_run_records_noalloc(storm_align_plan *plan, const int32_t *d_rec_rs, const int32_t *d_rec_diag,
                                    int64_t n_rec, storm_align_result *d_results, uint8_t *d_tb, void *stream);

/* ------------------------------------------------------------------------------------------
 * Seed index of one reference sequence, built on the device: index__build_reference
 * (index.c:19-195, declared index.h:49-62), with the contiguous layout of index.c:157-188.
 * packed:   the sequence as ReferenceDBType.sequence holds it (reference_data.h:25-53): 2-bit
 *           codes, code n in bits 2*(n&3) of byte n>>2 (codes.h:125-130), size codes.
 * region_start/region_end: the masked regions [start, end) (ReferenceDBType.masked_regions),
 *           ascending and disjoint; windows start-span .. end-2 are not indexed (index.c:66-72).
 * seed:     one seed pattern ('#' care, '-' don't care), as given to -s.
 * stdev_distance: -l; keys occurring more than avg + l*sd times are erased (index.c:117-156);
 *           <= 0 disables the filter.
 * storm_index_fetch copies first[4^w + 1] and map[n_map] (n_map = first[4^w]) to host buffers.
 * The mapper and the server build their indexes with the same kernels.
 */
typedef struct storm_index storm_index;
int storm_index_build(int device, const uint8_t *packed, int64_t size, const int32_t *region_start,
                      const int32_t *region_end, int n_regions, const char *seed, int stdev_distance,
                      storm_index **out);
int storm_index_sizes(const storm_index *ix, int64_t *n_keys, int64_t *n_map, int64_t *n_erased,
                      double *ms_device);
int storm_index_fetch(const storm_index *ix, int32_t *first, int32_t *map);
void storm_index_destroy(storm_index *ix);

/* ------------------------------------------------------------------------------------------
 * Group 2: whole mapping runs (reads_against_references, run.c:764-1109, with -M).
 * storm_map_files reads the same files as the reference (FASTA genome; FASTQ, FASTA or
 * csfasta(+.qual) reads), maps on the GPU and writes the reference's -M SAM.
 * mode: STORM_MODE_PLAIN runs the stages as separate launches per batch of equal-length reads;
 *       STORM_MODE_BATCHED runs every read of the batch, whatever its length, in one pass of
 *       persistent fused kernels.
 * argv/argc are echoed in the @PG line as the reference does (hit_map.c:310-315).
 */
#define STORM_MODE_PLAIN 0
#define STORM_MODE_BATCHED 1

typedef struct storm_map_stats {
  int64_t n_reads, n_mapped;
  int64_t n_candidates, n_filter_pass, n_accepted;
  int64_t n_lookups, n_hits;      /* seed-key lookups (first[key], first[key+1]) and raw bucket hits */
  int64_t n_fallback;             /* batched mode: reads the fused kernel handed to the staged kernels */
  double seconds_load, seconds_index, seconds_map, seconds_sam;
  double gpu_ms_seed, gpu_ms_filter, gpu_ms_align, gpu_ms_replay;
  double gpu_ms_lookup;           /* plain mode: the seed-key lookup kernel alone (part of gpu_ms_seed) */
} storm_map_stats;

int storm_map_files(int device, const char *genome_path, const char *reads_path, const char *qual_path,
                    const storm_params *params, int mode, const char *sam_path, int argc, char **argv,
                    storm_map_stats *stats);

/* In-memory variant: reads already packed (codes.h layout), one length per read.
 *   names        : n_reads NUL-terminated names (SAM QNAME)
 *   codes        : n_reads * code_stride bytes; qual: n_reads * qual_stride bytes of 2-bit
 *                  quality levels (read_quality.h:52-58) or NULL
 *   first_base / first_qual : color mode only (reads_data.h:22-27), else NULL
 * The reads are taken in the order given (the caller sorts them by tag if it wants the
 * reference's order, reads_data.c:59-67).  The SAM text is written to sam_path.  names may be
 * NULL: read i is then named "r<i>".
 */
typedef struct storm_reads_view {
  int64_t n_reads;
  const char *const *names;
  const uint8_t *codes;
  int64_t code_stride;
  const uint8_t *qual;
  int64_t qual_stride;
  const int32_t *len;
  const uint8_t *first_base;
  const uint8_t *first_qual;
  const char *db_name;   /* @RG ID/SM; NULL -> "*" */
  const uint8_t *no_qual; /* per read, 1 = the read has no quality (its FASTQ quality line is empty or
                           * starts with '#', which the reference reads as a comment,
                           * reads_data.c:459-532): QUAL '*' in the SAM, levels 3 in the DP (its qual
                           * row is 0xFF, first_qual 3); NULL = every read as the qual array says */
} storm_reads_view;

/* Reads loaded from a file into packed arrays ordered by tag: load_reads_db + sort_reads_db
 * (reads_data.h:63-68, reads_data.c:59-67,73-566).  use_device != 0 parses, packs and sorts FASTQ
 * files (four lines per record) on the device; FASTA, csfasta(+qual) and irregular FASTQ files are
 * loaded by the host.  storm_reads_get fills a view of the arrays (valid until
 * storm_reads_destroy), and optionally the n_reads tags, whether the device loaded them and its
 * kernel time.  storm_map_files loads FASTQ files this way. */
typedef struct storm_reads storm_reads;
int storm_reads_load(int device, const char *reads_path, const char *qual_path, const storm_params *params,
                     int use_device, storm_reads **out);
int storm_reads_get(const storm_reads *reads, storm_reads_view *view, uint64_t *tags, int *loaded_on_device,
                    double *device_ms);
void storm_reads_destroy(storm_reads *reads);

typedef struct storm_mapper storm_mapper;
/* Loads a FASTA genome, builds the seed indexes (host) and keeps genome + indexes resident in
 * HBM (index__build_reference, index.c:19-195; load_reference_db, reference_data.c:195-252). */
int storm_mapper_create(int device, const char *genome_path, const storm_params *params, storm_mapper **out);
int storm_mapper_map(storm_mapper *m, const storm_reads_view *reads, int mode, const char *sam_path,
                     int argc, char **argv, storm_map_stats *stats);
/* Per-read results without SAM text: for read r and rank k < 4, score/ref/start/sense
 * (HitInfoType, hit_map.h:17-40); unmapped slots have ref_id = -1. */
int storm_mapper_map_hits(storm_mapper *m, const storm_reads_view *reads, int mode,
                          int32_t *score, int32_t *ref_id, int32_t *ref_start, int32_t *sense,
                          storm_map_stats *stats);
/* The whole hit map of a mapping run in host memory, ready for SAM output: per read r and rank
 * k < 4 the entry hits[4r + k] (HitInfoType, hit_map.h:17-40; ref_id = -1 for an empty slot) and,
 * when tb is not NULL, the traceback of ranks k < max_hits (all four in the ordered mode) at
 * tb + (r * max_hits + k) * STORM_TB_BYTES in alignment__traceback's packing (TRACEBACK_SEQ_SET,
 * alignment.h:81-83; tb_len codes, the rest of the STORM_TB_BYTES zero, all zero for an empty slot).  This is what
 * hit_map__update leaves for hit_map__generate_SAM_output (hit_map.c:50-144, :197-504). */
typedef struct storm_hit {
  int32_t score, ref_id, ref_start, ref_end, read_start, read_end, sense, tb_len;
} storm_hit;
/* The tracebacks leave the device in the call's compact width (the longest read's codes plus the
 * indel cap) and are widened to STORM_TB_BYTES on the host (the environment variable
 * STORM_TB_FULL_WIDTH copies whole records instead); the bytes the caller receives are the same.
 * In the batched mode, on a one-sequence genome, a call may run as several fused launches over
 * consecutive chunks of the reads, the results of one chunk copied to hits/tb while the next runs
 * (the count from the copy volume and the band class's tuned launch time; the environment variable
 * STORM_RECORD_CHUNKS = K >= 1 fixes it).  The results are those of one launch. */
int storm_mapper_map_records(storm_mapper *m, const storm_reads_view *reads, int mode, storm_hit *hits, uint8_t *tb,
                             storm_map_stats *stats);
int64_t storm_mapper_free_bytes(storm_mapper *m);
void storm_mapper_destroy(storm_mapper *m);

/* ------------------------------------------------------------------------------------------
 * Group 2: serving loop.  No counterpart in the reference, which maps one read file per run
 * (reads_against_references, run.c:764-1109); each request here is such a run's -M mapping of
 * its reads (results per read as storm_mapper_map_hits returns them), and requests of different
 * sizes, read lengths, encodings and bands share the device.
 *
 * The genome (one FASTA file) and its seed indexes stay resident in HBM in the encodings asked
 * for (nucleotides for storm-nucleotide reads, colors for storm-color reads).  A request is
 * admitted when its reads and results fit the device-memory budget (a share of the free HBM at
 * creation, or a set budget); until then it waits, in arrival order.  Admitted requests are
 * mapped by continuous batching: each of `lanes` streams repeatedly takes up to max_batch_reads
 * unmapped reads of one parameter class (encoding, -i, -z, -t, scores, strands) from every
 * admitted request of that class, in admission order, and runs them through the persistent fused
 * kernel; lanes overlap on the device.  A read's result does not depend on the batch it lands in.
 * mode STORM_MODE_PLAIN is the per-request baseline: one request at a time through the staged
 * kernels, no admission control.  The index options (seeds, -l, -j) are the server's; a
 * request's params supply the rest.  submit and wait may be called from any thread.
 */
typedef struct storm_server storm_server;

typedef struct storm_server_config {
  int mode;                   /* STORM_MODE_BATCHED (default) or STORM_MODE_PLAIN */
  int lanes;                  /* batches in flight on the device, one stream each (default 3; plain: 1) */
  int cross_request;          /* 1 (default): a batch takes reads of several requests; 0: one request */
  int64_t max_batch_reads;    /* reads per batch (default 262144) */
  int64_t hbm_budget_bytes;   /* device memory for admitted requests; 0 = a quarter of the free HBM at
                                 creation; never more than free HBM - reserve */
  int64_t hbm_reserve_bytes;  /* left free for batch buffers and other users (default 8 GiB) */
  int encodings;              /* bit 0: nucleotide index, bit 1: color index (default 3) */
  const char *seeds_nucleotide;   /* NULL = reference defaults (seed.h:29) */
  const char *seeds_color;        /* NULL = reference defaults (seed.h:31) */
  int stdev_distance;         /* -l (default 100) */
  int ignore_lowercase;       /* -j */
  int sam_records;            /* 1 (default): each request also gets its -M SAM records (hit_map.c:325-504,
                                 formatted on the device with its batches); 0: hits only */
  int schedule;               /* which admitted request the next batch serves first:
                                 STORM_SCHED_SIZE (default): the earliest of submit time + 4 x the
                                 request's estimated remaining device time (read count x mean length^2
                                 x filter lanes, priced by the measured time of earlier batches), so
                                 a small request overtakes a large one and a large one is overtaken
                                 only by requests whose own estimate puts them first;
                                 STORM_SCHED_FIFO: admission order.  Results do not depend on it. */
} storm_server_config;
#define STORM_SCHED_FIFO 0
#define STORM_SCHED_SIZE 1

typedef struct storm_server_stats {
  int64_t submitted, completed, failed, reads_done;
  int64_t batches, batch_reads, batch_requests;   /* sums over batches */
  int64_t max_requests_per_batch;
  int64_t max_lanes_busy;       /* most batches in flight at once */
  int64_t admission_waits;      /* requests that waited for device memory */
  int64_t arena_bytes, arena_used_bytes, arena_peak_bytes;
  int64_t pending, active;      /* requests waiting for admission / admitted and unfinished */
  int64_t n_fallback, candidates;
  double gpu_busy_seconds;      /* summed over batches (lanes overlap, so it can exceed wall time) */
} storm_server_stats;

typedef struct storm_request_result {
  int64_t n_reads, n_mapped;
  int status;
  int batches;                  /* batches the request's reads were spread over */
  double seconds_queued;        /* submit -> admitted */
  double seconds_latency;       /* submit -> results ready */
  double t_submit, t_done;      /* seconds since the server was created */
} storm_request_result;

void storm_server_config_default(storm_server_config *cfg);
int storm_server_create(int device, const char *genome_path, const storm_server_config *cfg, storm_server **out);
/* Copies the reads (storm_reads_view: codes, qual, len, first_qual, and with sam_records names and
 * no_qual) and queues
 * them; *id identifies the request.  Fails at once when the request alone exceeds the budget. */
int storm_server_submit(storm_server *s, const storm_reads_view *reads, const storm_params *params, int64_t *id);
/* Blocks until request `id` is done, then copies its slots (n_reads x 4, as storm_mapper_map_hits)
 * into the non-NULL arrays and releases it.  Returns the request's status. */
int storm_server_wait(storm_server *s, int64_t id, int32_t *score, int32_t *ref_id, int32_t *ref_start, int32_t *sense,
                      storm_request_result *out);
/* As storm_server_wait, and the request's SAM records: the lines hit_map__generate_SAM_output writes
 * for its reads (hit_map.c:325-504, no header lines), in the request's read order, as
 * storm_mapper_map writes them for the same reads and parameters; names from the submitted view
 * (NULL names: "r<row>").  *sam is allocated by the library (release it with storm_free), NUL
 * terminated, *sam_bytes long.  The request's latency runs to the last record being ready.  Needs
 * sam_records = 1. */
int storm_server_wait_sam(storm_server *s, int64_t id, char **sam, int64_t *sam_bytes, int32_t *score, int32_t *ref_id,
                          int32_t *ref_start, int32_t *sense, storm_request_result *out);
void storm_free(void *p);
int storm_server_get_stats(storm_server *s, storm_server_stats *out);
/* Finishes every submitted request, then stops the lanes and frees the device memory. */
void storm_server_destroy(storm_server *s);

#ifdef __cplusplus
}
#endif
#endif /* STORM_GPU_H */
