// storm_kernels.h -- argument blocks and launchers of the gfx950 kernels (host and device).
#pragma once
#include "storm_dev.h"

namespace storm {

// one DP result (alignment__align + alignment__traceback outputs, alignment.h:189-229)
struct AlignRec {
  int32_t score;     // best score of the last row (MIN_SCORE if none)
  int32_t S;         // survival bound: min over rows >= 1 of max_on_line + best_possible
  int32_t win;       // window start (run.c:667)
  int16_t best_j, best_i0, best_j0, tb_len;
};

constexpr int TBB = (MAXL + MAX_I + 2) / 2;   // traceback bytes per record, 2 codes per byte

// one hit map entry (HitInfoType, hit_map.h:17-40); the traceback is (pass, rec)
struct Slot {
  int32_t score, ref_id, ref_start, ref_end;
  int32_t read_start, read_end, sense, pass, rec, tb_len;
  int32_t cell;   // traceback cell of this read (0..3) in the read's cell store (4 x TBB bytes per
                  // read); the four slots of a read always hold a permutation of the four cells
};
constexpr int PASS_CELL = -2;    // Slot.pass of entries whose traceback lives in the read's cell store

// one entry of the whole hit map in the caller's order (storm_hit of storm_gpu.h, HitInfoType)
struct HitRec {
  int32_t score, ref_id, ref_start, ref_end, read_start, read_end, sense, tb_len;
};

struct FilterArgs {
  GenomeDev g;
  const uint32_t* rplanes;   // read planes, candidate read index -> rplanes + idx * 2*RW
  const int32_t* rlen;       // read length, indexed by (read plane index >> len_shift)
  int len_shift;             // 1 when plane indices are read strands (2*read + strand)
  const int32_t* cand_read;  // read plane index per candidate
  const int32_t* cand_pos;   // window start per candidate
  int64_t n;
  uint8_t* score;            // optional u8 score per candidate
  uint8_t* pass;             // optional pass flag per candidate
  int mt, mm, op, ex, thr;
};

struct AlignArgs {
  GenomeDev g;
  ReadsDev r;
  ScoreDev sc;
  const int32_t* rec_rs;     // read-strand index (2*read + strand), global read numbering
  const int32_t* rec_diag;   // diagonal h
  int64_t n;                 // records in this launch
  int64_t base;              // index of the first record (into rec_* and outputs)
  uint8_t* scratch;          // direction words: ceil(n / 64) wavefronts x 64 x scratch_stride bytes
  int64_t scratch_stride;    // bytes per record: maxL x dp_dir_words(I) x 4
  AlignRec* out;
  uint8_t* tb;               // TBB bytes per record, traceback order (end to start)
  int check_bounds;          // 1: records given by a caller on the device are checked first (read
                             // strand inside r, window inside the padded genome); one outside is
                             // written as never accepted with best_j = -2 and not aligned
};

struct PrepArgs {
  const uint8_t* codes;      // n * code_stride bytes, 2-bit packed forward codes (codes.h:125-130)
  int code_stride;
  const uint8_t* qual;       // n * qual_stride bytes, 2-bit packed levels, or null
  int qual_stride;
  const uint8_t* first_qual; // color: level of the primer-side color (reads_data.h:27), or null
  const int32_t* len;
  int64_t n;
  int color;
  uint32_t* planes;          // out: (2n) * 2RW
  uint32_t* qplanes;         // out: (2n) * QW
};

struct SeedArgs {
  ReadsDev r;
  SeedSetDev seeds;
  int I;
  int sense_mask;            // bit 0 forward, bit 1 reverse
  int64_t rs_begin;          // first read strand of this launch (2*read + strand)
  int64_t n_rs;
  int64_t* seg_count;        // hits (pass 1) then kept candidates (pass 2) per read strand
  const int64_t* seg_off;    // exclusive scan of the pass-1 counts (n_rs + 1)
  int32_t* hits;             // raw diagonals of overflow segments at seg_off
  int32_t* kept;             // kept candidates of each segment at seg_off
  uint8_t* overflow;         // 1 when the segment did not fit in LDS
  int32_t* n_overflow;       // count of such segments
};

struct CompactArgs {
  const int64_t* seg_off;
  const int64_t* seg_count;
  const int64_t* cand_off;   // exclusive scan of seg_count (n_rs + 1)
  const int32_t* kept;
  int64_t rs_begin, n_rs;
  int d;
  int32_t* cand_rs;
  int32_t* cand_diag;
  int32_t* cand_pos;         // diag - d (filter window start, run.c:651)
};

struct RecArgs {
  const uint8_t* pass;
  const int64_t* pass_scan;  // exclusive scan of pass flags (n_cand + 1 entries)
  const int32_t* cand_rs;
  const int32_t* cand_diag;
  int64_t n_cand;
  const int64_t* cand_off;   // n_rs + 1
  int64_t n_rs;
  int32_t* rec_rs;
  int32_t* rec_diag;
  int64_t* rec_off;          // n_rs + 1
};

struct ReplayArgs {
  const AlignRec* rec;
  const int64_t* rec_off;    // per read strand of the launch, n_rs + 1 entries
  const int32_t* len;
  int64_t read_begin;        // first read of the launch
  int64_t n_reads;
  int ref_id;
  int pass;
  int t;
  int sense_mask;
  Slot* slots;               // 4 per read, global read id
};

// persistent fused kernel (k_fused.hip)
struct FusedArgs {
  GenomeDev g;
  ReadsDev r;
  SeedSetDev seeds;
  ScoreDev sc;
  const int32_t* n_reads;    // {end, first}: the launch maps reads [first, end) (device values, so a
                             // captured graph serves any size and any range of the batch)
  int group;                 // G reads per task
  int32_t* task_counter;     // cleared before the launch
  Slot* slots;               // 4 per read
  uint8_t* cells;            // 4 * TBB bytes per read: the traceback cells
  uint8_t* overflow;         // per read
  uint8_t* scratch;          // per workgroup: 64 * scratch_stride bytes of direction words
  int64_t scratch_stride;    // bytes per lane: maxL x dp_dir_words(I) x 4
  uint8_t* tbbuf;            // per workgroup: FUSED_PCAP * TBB traceback bytes of the current flush
  uint8_t* work;             // per workgroup: FUSED_WORK_BYTES of candidates and records
  int ref_id;
  int sense_mask;
  unsigned long long* stats; // [0] candidates, [1] filter passes, [2] accepted, [3] raw hits
  unsigned long long* phase; // measurement build only (STORM_FUSED_PHASES): wavefront clock ticks
                             // per phase, FUSED_PHASES entries; null otherwise
};
// phases of k_fused: candidates (seed keys, lookups, gather, sort, keep rule), filter, pass scan +
// DP triage, DP + traceback, hit-map replay + traceback cells, task fetch
constexpr int FUSED_PHASES = 6;

constexpr int FUSED_PCAP = 1024;   // DP records per flush of k_fused
constexpr int FUSED_CCAP = 8192;   // candidates per flush of k_fused
// per-workgroup work buffer of k_fused: diagonal + record index (int32) per candidate, AlignRec
// + candidate index per record, read strand + filter flag (u8) per candidate
constexpr int64_t FUSED_WORK_BYTES =
    ((8LL * FUSED_CCAP + ((int64_t)sizeof(AlignRec) + 4) * FUSED_PCAP + 2LL * FUSED_CCAP) + 255) & ~255LL;

// serving loop (k_serve.hip): one admitted request's reads and results in the device arena
struct ServeRegion {
  const uint8_t* codes;      // n * code_stride bytes, 2-bit packed (codes.h:125-130)
  const uint8_t* qual;       // n * qual_stride bytes of 2-bit levels, or null (every level 3)
  const int32_t* len;
  const uint8_t* first_qual; // color: level of the primer-side color, or null
  int32_t code_stride, qual_stride;
  int32_t* hits;             // out: per read, 4 slots x {score, ref_id, ref_start, sense}
};
struct ServeRow {            // one read of a batch: region (index into the batch's table), read
  int32_t region, row;
};
constexpr int SERVE_CS = MAXL / 4;   // code and quality bytes per read in a batch
struct ServeGatherArgs {
  const ServeRegion* regions;
  const ServeRow* rows;
  int64_t n;
  uint8_t* codes;            // n * SERVE_CS
  uint8_t* qual;             // n * SERVE_CS
  int32_t* len;
  uint8_t* first_qual;
};
hipError_t launch_serve_gather(const ServeGatherArgs& a, hipStream_t st);
hipError_t launch_serve_scatter(const Slot* slots, const ServeRow* rows, int64_t n, const ServeRegion* regions,
                                hipStream_t st);

// The DP kernels are compiled for band bounds IM = 1, 3, 7, 15, 31 (2IM + 1 columns); -i I uses the
// smallest IM >= I.  Traceback directions are 4 bits per cell, 8 cells per 32-bit word.
__host__ __device__ inline int dp_band_bound(int I) { return I <= 1 ? 1 : I <= 3 ? 3 : I <= 7 ? 7 : I <= 15 ? 15 : 31; }
// records whose direction words one wavefront of k_align interleaves
constexpr int ALIGN_WAVE_RECORDS = 64;
__host__ __device__ inline int dp_dir_words(int I) { return (2 * dp_band_bound(I) + 1 + 7) / 8; }

// FASTQ reads parsed on the device (k_ingest.hip)
struct IngestArgs {
  const uint8_t* text;       // the file's records (after the leading comment lines)
  int64_t n_bytes;
  const int64_t* nl;         // byte offsets of the newlines
  int64_t n_nl;
  int64_t n_reads;           // records of four lines
  int color, qual_base;
  int32_t* len;              // out: read length, -1 irregular record, -2 empty read
  int64_t* header;           // out: offset of the header line
  uint8_t* hasq;             // out: quality line not empty
  uint8_t* codes;            // out: n_reads x code_stride packed codes (zeroed)
  uint8_t* qual;             // out: packed quality levels (zeroed), or null
  int64_t code_stride;
  uint64_t* tag;             // out
  uint8_t* first_base;       // out (color)
  uint8_t* first_qual;       // out (color)
  uint8_t* err;              // out: quality character below the base
};
int64_t nl_chunks(int64_t n_bytes);
hipError_t launch_nl_count(const uint8_t* d, int64_t n, int32_t* counts, int64_t n_chunks, hipStream_t st);
hipError_t launch_nl_positions(const uint8_t* d, int64_t n, const int32_t* first, int64_t n_chunks, int64_t* nl,
                               hipStream_t st);
hipError_t launch_fastq_len(const IngestArgs& a, hipStream_t st);
hipError_t launch_fastq_pack(const IngestArgs& a, hipStream_t st);
hipError_t launch_iota_i32(int32_t* v, int64_t n, hipStream_t st);
hipError_t sort_tags(uint64_t* k0, uint64_t* k1, int32_t* v0, int32_t* v1, int64_t n, void* tmp, size_t& tmp_bytes,
                     int& result_sel, hipStream_t st);

// genome planes and the device index build (k_index.hip)
hipError_t launch_unpack_codes(const uint8_t* packed, int64_t size, uint8_t* code, hipStream_t st);
hipError_t launch_genome_planes(const uint8_t* code, const uint8_t* mask, int64_t size, int64_t nw, uint2* lh,
                                uint32_t* mk, hipStream_t st);
hipError_t launch_index_keys(const uint8_t* code, int64_t map_size, const int2* regions, int n_regions,
                             const SeedDev& seed, uint32_t ks, uint32_t* keys, int32_t* vals, hipStream_t st);
hipError_t launch_index_runs(const uint32_t* keys, int64_t n, int32_t* st_, int32_t* en, hipStream_t st);
hipError_t launch_index_counts(const int32_t* st_, int32_t* en, int64_t ks, hipStream_t st);
hipError_t launch_index_erase(const uint8_t* erased, int32_t* counts, int64_t ks, hipStream_t st);
hipError_t launch_index_scatter(const uint32_t* keys, const int32_t* vals, int64_t n, const int32_t* st_,
                                const int32_t* first, const uint8_t* erased, uint32_t ks, int32_t* map,
                                hipStream_t st);
hipError_t sort_key_pos_raw(uint32_t* k0, uint32_t* k1, int32_t* v0, int32_t* v1, int64_t n, int end_bit, void* tmp,
                            size_t& tmp_bytes, int& result_sel, hipStream_t st);
hipError_t scan_exclusive_i32(const int32_t* in, int32_t* out, int64_t n, void* tmp, size_t& tmp_bytes,
                              hipStream_t st);
hipError_t launch_fused(const FusedArgs& a, int grid, hipStream_t st);
hipError_t launch_fused_reset(int32_t* counter, unsigned long long* stats, hipStream_t st);
int fused_blocks_per_cu(int W, int I, int color);
hipError_t launch_filter(const FilterArgs& a, int W, hipStream_t st);
hipError_t launch_align(const AlignArgs& a, hipStream_t st);
hipError_t launch_prep_reads(const PrepArgs& a, hipStream_t st);
hipError_t launch_seed_count(const SeedArgs& a, hipStream_t st);
hipError_t launch_seed_gather(const SeedArgs& a, hipStream_t st);
hipError_t launch_unique_segments(const SeedArgs& a, hipStream_t st);
hipError_t launch_compact_candidates(const CompactArgs& a, hipStream_t st);
hipError_t launch_make_records(const RecArgs& a, hipStream_t st);
hipError_t launch_replay(const ReplayArgs& a, hipStream_t st);
hipError_t launch_slots_init(Slot* slots, int64_t n, hipStream_t st);
// SAM records of the -M output formatted on the device (k_sam.hip), caller rows [row0, row0 + n)
struct SamArgs {
  const Slot* slots;          // device order, 4 per read
  const uint8_t* cells;       // traceback cells, 4 x TBB bytes per read (device order)
  const int32_t* inv;         // caller row -> device row, or null (the same order)
  int64_t row0, n;
  const uint8_t* codes;       // device order: packed forward codes (codes.h:125-130)
  int64_t code_stride;
  const uint8_t* qual;        // device order: packed 2-bit levels, or null
  int64_t qual_stride;
  const uint8_t* first_qual;  // color: level of the primer-side color (device order), or null
  const int32_t* len;         // device order
  const uint8_t* no_qual;     // caller order: 1 = the read has no quality, or null
  const char* names;          // caller order: row r's name at names[name_off[r] .. name_off[r + 1]),
  const int64_t* name_off;    // or null: "r<name_base + r>"
  int64_t name_base;
  const GenomeDev* genomes;   // per reference sequence
  const uint8_t* sym_ck;      // color: per sequence, the base symbol before each 32-code word
  const int64_t* sym_off;     // (hit_map.c:236-249), sequence r's at sym_ck + sym_off[r]
  const char* ref_names;      // RNAME of sequence r at ref_names[ref_name_off[r] .. ref_name_off[r + 1])
  const int32_t* ref_name_off;
  const uint8_t* mapq_lut;    // MAPQ by score difference + 32768 (65536 entries)
  int M, uniqueness, print_all, qual_base, color, t;
  int64_t* lens;              // count pass: bytes of row row0 + v
  int64_t* mapped;            // count pass: reads with a hit (added to), or null
  const int64_t* offs;        // write pass: where row row0 + v starts in text
  char* text;
};
hipError_t launch_sam(const SamArgs& a, bool write, hipStream_t st);

// tracebacks of the entries pass `pass` left in ranks < max_rank of slots [slot_begin, slot_begin +
// n_slots), from the pass's records (tb, TBB bytes per record) into the read's cell store at
// cells + (id * 4 + cell) * TBB, id = ids ? ids[read] : read; those entries become PASS_CELL
hipError_t launch_store_cells(Slot* slots, int64_t slot_begin, int64_t n_slots, int pass, int max_rank, const uint8_t* tb,
                              uint8_t* cells, const int32_t* ids, hipStream_t st);
// the slots of device rows [r0, r1) as HitRec rows order[i] (or i) of out, and the tracebacks of
// ranks < max_rank into tb_out + (row * max_rank + rank) * tb_stride (tb_out may be null): the first
// tb_stride bytes of each (a multiple of 16, at most TBB, at least the longest traceback's bytes);
// *mapped (may be null) is raised by the rows whose first slot holds a hit
hipError_t launch_hit_records(const Slot* slots, const int32_t* order, int64_t r0, int64_t r1, int max_rank,
                              const uint8_t* cells, HitRec* out, uint8_t* tb_out, int tb_stride, int* mapped,
                              hipStream_t st);
hipError_t launch_subset_gather(const uint32_t* planes, const uint32_t* qplanes, const int32_t* len, const Slot* slots,
                                const int32_t* ids, int64_t k, uint32_t* sp, uint32_t* sq, int32_t* sl, Slot* ss,
                                hipStream_t st);
hipError_t launch_subset_scatter(const Slot* ss, const int32_t* ids, int64_t k, Slot* slots, hipStream_t st);
hipError_t launch_permute_bytes(const uint8_t* src, int64_t stride, const int32_t* order, int64_t n, uint8_t* dst,
                                hipStream_t st);
hipError_t launch_permute_i32(const int32_t* src, const int32_t* order, int64_t n, int32_t* dst, hipStream_t st);
hipError_t launch_hits_out(const Slot* slots, const int32_t* order, int64_t n, int32_t* score, int32_t* ref_id,
                           int32_t* ref_start, int32_t* sense, hipStream_t st);
hipError_t launch_overflow_ranges(const int64_t* seg_off, const uint8_t* overflow, int64_t n_rs, int64_t* begin,
                                  int64_t* end, hipStream_t st);
// device-wide helpers (rocPRIM)
hipError_t scan_exclusive_i64(const int64_t* in, int64_t* out, int64_t n, void* tmp, size_t& tmp_bytes, hipStream_t st);
hipError_t scan_exclusive_u8(const uint8_t* in, int64_t* out, int64_t n, void* tmp, size_t& tmp_bytes, hipStream_t st);
hipError_t segsort_i32(const int32_t* in, int32_t* out, int64_t n_items, int64_t n_seg, const int64_t* begin,
                       const int64_t* end, void* tmp, size_t& tmp_bytes, hipStream_t st);

}  // namespace storm
