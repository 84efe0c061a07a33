// mapper.h -- internal interface of the mapping pipeline, shared by the C ABI (capi.cpp) and the
// serving loop (server.cpp): device buffers, the resident genome planes and seed indexes, and the
// stages of one mapping batch.
#pragma once
#include <hip/hip_runtime.h>

#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "host_io.h"
#include "storm_kernels.h"

namespace storm {

// sets the calling thread's storm_last_error() message and returns `code`
int fail(int code, const char* fmt, ...);

#define HIPOK(expr)                                                                         \
  do {                                                                                      \
    hipError_t _e = (expr);                                                                 \
    if (_e != hipSuccess) return ::storm::fail(STORM_ERR_DEVICE, "%s: %s (%s:%d)", #expr,   \
                                               hipGetErrorString(_e), __FILE__, __LINE__);  \
  } while (0)

// device allocation owned by a C++ object
struct DevMem {
  void* p = nullptr;
  size_t n = 0;
  DevMem() = default;
  DevMem(const DevMem&) = delete;
  DevMem& operator=(const DevMem&) = delete;
  DevMem(DevMem&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevMem& operator=(DevMem&& o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevMem() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t ensure(size_t bytes) {
    if (bytes <= n && p) return hipSuccess;
    release();
    size_t b = bytes ? bytes : 16;
    hipError_t e = hipMalloc(&p, b);
    if (e == hipSuccess) n = b;
    return e;
  }
  template <class T> T* as() const { return (T*)p; }
};

double now_s();

// genome bit planes of one reference sequence
struct DevGenome {
  DevMem lh, mask;   // {lo, hi} word pairs and the mask plane
  int64_t size = 0;
  GenomeDev view() const { return GenomeDev{lh.as<uint2>(), mask.as<uint32_t>(), size}; }
};

struct DevIndex {
  DevMem first, map;
  SeedSpec spec;
  int64_t n_map = 0;   // positions held (first[4^w])
};

// scratch of the device index build, reused across the seeds of a sequence
struct IndexWork {
  DevMem keys0, keys1, vals0, vals1, st, en, erased, tmp;
  int32_t* host_counts = nullptr;   // pinned, 4^w key counts for the repeat filter
  int64_t host_counts_n = 0;
  std::vector<uint8_t> host_erased;
  IndexWork() = default;
  IndexWork(const IndexWork&) = delete;
  IndexWork& operator=(const IndexWork&) = delete;
  ~IndexWork();
};

struct RefState {
  DevGenome g;
  std::vector<DevIndex> idx;
  bool resident = false;
};

// Genome (one encoding: nucleotides or colors) and its seed indexes, shared by every mapper that
// maps against it.  With `pinned` every sequence stays resident; otherwise a multi-sequence genome
// is uploaded one sequence at a time (each pass releases the previous one).
// Tables the SAM formatter reads, built once per resident genome: the sequences' planes, their
// names, and in color space the base symbol before every 32-code word of each sequence.
struct SamTables {
  std::mutex mu;
  bool ready = false;
  DevMem genomes, sym_ck, sym_off, names, name_off;
};

struct Resident {
  std::vector<HostRef> refs;
  std::vector<SeedSpec> seeds;
  std::string seeds_str;
  std::vector<std::unique_ptr<RefState>> rstate;
  double index_seconds = 0;
  bool pinned = false;
  SamTables sam;
};

// pinned host memory owned by a C++ object
struct PinnedMem {
  void* p = nullptr;
  size_t n = 0;
  PinnedMem() = default;
  PinnedMem(const PinnedMem&) = delete;
  PinnedMem& operator=(const PinnedMem&) = delete;
  ~PinnedMem() { release(); }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t ensure(size_t bytes) {
    if (bytes <= n && p) return hipSuccess;
    release();
    const size_t b = bytes ? bytes : 16;
    hipError_t e = hipHostMalloc(&p, b, hipHostMallocDefault);
    if (e == hipSuccess) n = b;
    return e;
  }
};

struct PassScratch {   // reused across passes
  DevMem seg_count, seg_off, hits, kept, overflow, n_overflow, seg_begin, seg_end, sorted;
  DevMem cand_off, cand_rs, cand_diag, cand_pos, pass, pass_scan;
  DevMem rec_rs, rec_diag, rec_off, recs, tb, align_scratch, cub_tmp;
};

// reads resident on the device for one map call (global read ids = device order)
struct DevReads {
  DevMem codes, qual, first_qual, len, planes, qplanes;
  int64_t n = 0;
  int has_qual = 0;
  ReadsDev view() const {
    return ReadsDev{planes.as<uint32_t>(), qplanes.as<uint32_t>(), len.as<int32_t>(), n, has_qual};
  }
};

// batched path: buffers sized to the largest batch seen, one captured graph per reference
struct FusedTune {
  int group = 16;
  int grid = 0;
  float ms = 0;
};

struct FusedState {
  DevReads reads;
  DevMem slots, cells, overflow, n_dev, counter, stats, scratch, tbbuf, work, gids, gtb, gcount;
  DevMem phase;   // measurement build only: k_fused's per-phase clock ticks
  int64_t cap_reads = 0, scratch_stride = 0;
  int grid = 0;
  // captured graphs (reset + fused kernel) keyed by the buffers, settings and reference they were
  // captured with; a context that alternates parameter classes keeps one per class
  std::vector<std::pair<std::vector<const void*>, hipGraphExec_t>> graphs;
  static constexpr size_t MAX_GRAPHS = 16;
  std::map<int64_t, FusedTune> tune;    // key: (W, I, color, length bucket)
  bool tuned_any = false;
  ~FusedState() { drop_graphs(); }
  hipGraphExec_t find_graph(const std::vector<const void*>& sig) const {
    for (const auto& g : graphs)
      if (g.first == sig) return g.second;
    return nullptr;
  }
  void add_graph(const std::vector<const void*>& sig, hipGraphExec_t e) {
    if (graphs.size() >= MAX_GRAPHS) {
      (void)hipGraphExecDestroy(graphs.front().second);
      graphs.erase(graphs.begin());
    }
    graphs.emplace_back(sig, e);
  }
  void drop_graph(const std::vector<const void*>& sig) {
    for (size_t k = 0; k < graphs.size(); ++k)
      if (graphs[k].first == sig) {
        (void)hipGraphExecDestroy(graphs[k].second);
        graphs.erase(graphs.begin() + (long)k);
        return;
      }
  }
  void drop_graphs() {
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
    graphs.clear();
  }
};

// autotuner results shared by the contexts of one device (the serving loop's lanes)
struct SharedTune {
  std::mutex mu;
  std::map<int64_t, FusedTune> best;
};

struct PassStats {
  int64_t cand = 0, rec = 0, hits = 0;
  float ms_seed = 0, ms_filter = 0, ms_align = 0, ms_replay = 0, ms_lookup = 0;
};

// where a staged pass leaves the tracebacks it inserts: the cell store of the reads (4 x TBB bytes
// per read), read r's cells at (ids ? ids[r] : r) * 4
struct CellTarget {
  uint8_t* cells;
  const int32_t* ids;
};

}  // namespace storm

// One execution context: a stream, per-batch buffers and the parameters of the batches it maps,
// over a shared resident genome.
struct storm_mapper {
  int device = 0;
  hipStream_t st = nullptr;
  storm_params p{};
  std::string unmapped_str;    // the -O path p.unmapped_path points to (empty: none)
  std::shared_ptr<storm::Resident> res;
  storm::PassScratch ps;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;   // the fused launch's timing events (made on first use)
  storm::FusedState fs;
  storm::DevReads stage;       // the caller's reads as uploaded, before the device reordering
  storm::DevReads plain_reads; // the plain path's reads in device order, reused across calls
  storm::DevMem plain_slots;   // the plain path's hit map slots, reused across calls
  storm::DevMem plain_cells;   // the plain path's traceback cells (4 per read), reused across calls
  std::vector<int64_t> len_hist;   // reads per length of the current call (length_order_chunks)
  storm::DevMem rec_dev, tb_dev;   // the hit map and tracebacks in the caller's order, on their way out
  // SAM records formatted on the device: the caller's read names, the MAPQ table of the scores,
  // caller row -> device row, per-row byte counts and offsets, the text and its host copy
  storm::DevMem sam_names, sam_name_off, sam_no_qual, sam_lut, sam_inv, sam_lens, sam_offs, sam_text, sam_mapped;
  int sam_lut_dm = 0;
  storm::PinnedMem sam_host;
  storm::PinnedMem out_stage;          // two pinned halves the results pass through to the caller
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
  // chunked records calls: the stream results leave on while the next chunk runs on st, the event
  // that orders them after the chunk's records kernel, and the chunks' {end, first} read ranges
  hipStream_t st_out = nullptr;
  hipEvent_t rec_ev = nullptr;
  storm::PinnedMem ranges;
  double finish_seconds = 0;   // time spent in a map call's finish step (SAM formatting)
  storm::DevReads sub_reads;   // the batched path's staged fallback: the handed-back reads' planes,
  storm::DevMem sub_ids, sub_slots;   // their ids in the batch and their slots
  storm::DevMem order_dev;     // device row -> caller's read index
  int64_t order_dev_n = -1;    // reads of the current call whose order order_dev holds (-1: none)
  std::vector<int32_t> order32;   // the current call's order in 32 bits, on its way to order_dev
  // host arrays of a call, kept so that a call of a size seen before touches no fresh pages: the
  // device order, the fused kernel's per-read overflow flags
  std::vector<int64_t> order64;
  std::vector<uint8_t> ovf_host;
  storm::DevMem mapped_dev;    // reads with a hit, counted by the records kernel
  storm::DevMem hits_dev;      // per-slot results on their way to the caller
  storm::SharedTune* shared_tune = nullptr;   // tuner results of sibling contexts, or null
  bool want_tb = true;   // false while a call returns hits only: tracebacks stay on the device
  storm_mapper() = default;
  storm_mapper(const storm_mapper&) = delete;
  storm_mapper& operator=(const storm_mapper&) = delete;
  ~storm_mapper() {
    for (hipEvent_t e : {ev0, ev1, stage_ev[0], stage_ev[1], rec_ev})
      if (e) (void)hipEventDestroy(e);
    if (st_out) (void)hipStreamDestroy(st_out);
  }
};

namespace storm {

int filter_lanes(int d);
void fill_score(const storm_params& p, ScoreDev& sc);
// hit map ranks whose tracebacks the SAM needs: the -M reported ranks, all four in ordered mode
int tb_ranks(const storm_params& p);
const char* default_seeds(int color);
int check_params(const storm_params* p);

// Codes (one per byte) and mask bytes (or null) of a sequence to the device bit planes; with
// keep_code the device copy of the codes is handed over (for the index build).
int upload_genome(const uint8_t* code, const uint8_t* mask, int64_t size, DevGenome& g, hipStream_t st,
                  DevMem* keep_code = nullptr);
// index__build_reference (index.c:19-195) on the device: d_code = the sequence's codes, one per
// byte; d_regions = its masked regions [start, end), sorted.  ix.first / ix.map in the contiguous
// layout (index.c:157-188).
int build_index_dev(const uint8_t* d_code, int64_t size, const int2* d_regions, int n_regions, const SeedSpec& seed,
                    int stdev_distance, DevIndex& ix, IndexWork& w, hipStream_t st, int64_t* erased_out);
// every seed's index of one sequence; erased (optional) receives the erased key count per seed
int build_indexes(const HostRef& ref, const DevMem& d_code, const std::vector<SeedSpec>& seeds, int stdev_distance,
                  std::vector<DevIndex>& out, hipStream_t st, std::vector<int64_t>* erased = nullptr);

// FASTQ reads parsed, packed and ordered by tag on the device (read_ingest.cpp).  handled = false
// when the file is not a four-lines-per-record FASTQ file: the caller then uses load_reads.
int load_reads_device(int device, const char* path, const storm_params& p, HostReads& out, bool& handled,
                      double* device_ms);

// Loads a FASTA genome in the encoding of p.color and parses the seeds (no device work).
int load_resident(const char* genome_path, const storm_params& p, std::shared_ptr<Resident>& out);
int ensure_ref_resident(storm_mapper* m, size_t r);

// Packed reads already in d.codes / d.qual / d.first_qual / d.len (device): builds the strand planes.
// DP + traceback (k_align) of records [0, n_rec) of `aa`, maxL = the longest read among them: launches
// over chunks of whole wavefronts sized to a 256 MiB direction scratch, grown in `scratch`
int align_records(AlignArgs aa, int maxL, int64_t n_rec, DevMem& scratch, hipStream_t st);
// the resident inputs of a storm_align_plan (its kernel arguments with the plan's own records and
// outputs), the longest read, its device and stream
int align_plan_view(storm_align_plan* pl, AlignArgs& a, int& maxL, int& device, hipStream_t& st);
// bytes of direction scratch align_records needs for n_rec records of reads up to maxL at -i I
int64_t align_scratch_bytes(int I, int maxL, int64_t n_rec);
int prep_device_reads(storm_mapper* m, int64_t n, int64_t code_stride, int64_t qual_stride, bool has_qual,
                      bool has_first_qual, DevReads& d);
// order32: the order in 32 bits, when the caller has it.  A non-identity order is left in
// m->order_dev (m->order_dev_n = the view's reads) for the call's result copies.
int upload_reads(storm_mapper* m, const storm_reads_view* v, const std::vector<int64_t>& order, DevReads& d,
                 const std::vector<int64_t>* hist = nullptr, const bool* identity = nullptr,
                 const std::vector<int32_t>* order32 = nullptr);
// one staged pass over reads [r0, r1) of dr; PASS_SPLIT (nothing written) when its raw index hits
// exceed PASS_HIT_LIMIT
constexpr int PASS_SPLIT = 1;
constexpr int64_t PASS_HIT_LIMIT = 1LL << 30;
int run_pass(storm_mapper* m, DevReads& dr, Slot* slots, const CellTarget& ct, int ref_id, int pass_id, int64_t r0,
             int64_t r1, PassStats& stats);
// reads [r0, r1) in passes of at most `chunk` reads, split further as PASS_SPLIT asks
int run_passes(storm_mapper* m, DevReads& dr, Slot* slots, const CellTarget& ct, int ref_id, int& pass_id, int64_t r0,
               int64_t r1, PassStats& stats, int64_t chunk);
// the staged kernels over reads dev_ids of m->fs.reads, slots read from and written back to m->fs.slots
int staged_subset(storm_mapper* m, const std::vector<int64_t>& dev_ids, int ref, int& pass_id, PassStats& ps);

// Reads that the fused kernel handed back (device ids of m->fs.reads) for reference `ref`; the
// callback maps them with the staged kernels and leaves their slots in m->fs.slots.
using FallbackFn = std::function<int(const std::vector<int64_t>& dev_ids, int ref, int& pass_id, PassStats& ps)>;
// The batched path over the n reads already in m->fs.reads (device order, longest maxL): slots
// initialised, then per reference the captured fused kernel and the fallback of its overflow
// reads.  Slots end in m->fs.slots.  totals: candidates, filter passes, accepted, raw hits.
int fused_core(storm_mapper* m, int64_t n, int maxL, const FallbackFn& fallback, PassStats& ps,
               unsigned long long totals[4], int64_t& n_fallback);

// caller arrays for the per-read slots (n_reads x 4 each, any may be null)
struct HitsOut {
  int32_t *score, *ref_id, *ref_start, *sense;
};
// the state a map call leaves on the device for its output: slots (device order, 4 per read), the
// reads' traceback cells, the reads (device order) and device row -> caller row
struct FinishCtx {
  const Slot* slots;
  const uint8_t* cells;
  const DevReads* reads;
  const std::vector<int64_t>* order;
  int64_t n;
};
// replaces the default finish (records to host) when given; sets *mapped
using FinishFn = std::function<int(const FinishCtx& fc, int64_t& mapped)>;
// the planes of every sequence resident (the SAM formatter reads them all)
int ensure_planes(storm_mapper* m);
// The rows a SAM formatting pass covers, in output order (row r = caller row r of the map call):
// names packed (row r's at names[name_off[r] .. name_off[r + 1]); names = null: "r<r>"), which
// rows have no quality (QUAL '*'), and the strides of the device copies of the reads.
struct SamRows {
  int64_t n;
  const char* names;
  const int64_t* name_off;
  const uint8_t* no_qual;     // per row, or null
  bool has_qual, has_first_qual;
  int64_t code_stride, qual_stride;
};
// receives the text of rows [row0, row0 + n) and, when the pass asks for them, the rows' start
// offsets in it (n + 1 entries, relative to text)
using SamSink = std::function<int(const char* text, int64_t bytes, int64_t row0, int64_t n, const int64_t* row_off)>;
// the -M SAM records of a finished map call formatted on the device, handed to `sink` in chunks of
// rows in row order; with_offsets: the sink also gets each row's offset
int sam_device(storm_mapper* m, const SamRows& rows, const FinishCtx& fc, const SamSink& sink, bool with_offsets,
               int64_t& mapped);
// SamRows of a storm_reads_view (names packed into `names` / `name_off`)
SamRows sam_rows_of_view(const storm_reads_view* v, std::string& names, std::vector<int64_t>& name_off);

// the default finish: slots to the caller's arrays (ho), whole records (ro), or host slots
struct RecordsOut;
int finish_view(storm_mapper* m, const Slot* slots, const uint8_t* cells, const std::vector<int64_t>& order, int64_t n,
                const HitsOut* ho, const RecordsOut* ro, std::vector<HostSlot>& out_slots, int64_t& mapped);

// caller arrays for whole hit records: hits[4 * read + rank], tracebacks of ranks < tb_ranks(p)
// at tb + (read * tb_ranks(p) + rank) * TBB (tb may be null)
struct RecordsOut {
  storm_hit* hits;
  uint8_t* tb;
};
// ho != null: the slots go straight to the caller's arrays (no host slot assembly, no tracebacks);
// ro != null: the records and tracebacks go to the caller's arrays instead of out_slots
int map_view(storm_mapper* m, const storm_reads_view* v, int mode, std::vector<HostSlot>& out_slots,
             storm_map_stats* st_out, const HitsOut* ho = nullptr, const RecordsOut* ro = nullptr,
             const FinishFn* fin = nullptr);

}  // namespace storm
