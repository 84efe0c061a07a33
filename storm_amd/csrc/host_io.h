// host_io.h -- host-side data of the mapping path: genome, reads, seeds, CSR index, SAM text.
#pragma once
#include <cstdint>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

#include "../../include/storm_gpu.h"

namespace storm {

struct HostRef {                       // ReferenceDBType (reference_data.h:25-53)
  std::string name;                    // empty -> "*"
  int64_t size = 0;                    // codes (colors in color mode)
  uint8_t first_base = 0xff;           // color mode
  std::vector<uint8_t> code;           // one code per byte
  std::vector<uint8_t> mask;           // 1 = masked
  std::vector<std::pair<int, int>> regions;   // masked regions [start, end)
};

struct HostReads {                     // ReadsDBType (reads_data.h:46-55), one length per read
  std::string db_name;                 // empty -> "*"
  int64_t n = 0;
  int max_len = 0;
  bool has_qual = false;
  int64_t code_stride = 0, qual_stride = 0;
  std::vector<std::string> names;
  std::vector<int32_t> len;
  std::vector<uint8_t> codes;          // 2-bit packed, code_stride bytes per read
  std::vector<uint8_t> qual;           // 2-bit quality levels, qual_stride bytes per read
  std::vector<uint8_t> first_base, first_qual;   // color mode
  std::vector<uint64_t> tag;
  std::vector<uint8_t> no_qual;        // per read (or empty): quality line ignored, QUAL '*'
};

struct SeedSpec {                      // SeedType (seed.h:41-55)
  int weight = 0, span = 0;
  std::vector<int> offs, pos;
};

struct HostIndex {                     // IndexType (index.h:22-41), contiguous layout
  std::vector<int32_t> first, map;
  int64_t erased = 0;
};

int parse_seeds(const char* src, std::vector<SeedSpec>& out, std::string& err);
int load_reference(const char* path, const storm_params& p, std::vector<HostRef>& out, std::string& err);
int load_reads(const char* path, const char* qual_path, const storm_params& p, HostReads& out, std::string& err);
void sort_reads_by_tag(HostReads& r);
void build_index(const HostRef& ref, const SeedSpec& seed, const storm_params& p, HostIndex& out);
int64_t repeat_filter(const int32_t* counts, int64_t ksize, int64_t count, int stdev_distance, uint8_t* erased);

struct HostSlot {
  int32_t score, ref_id, ref_start, ref_end, read_start, read_end, sense;
  std::vector<uint8_t> tb;   // traceback in alignment order (end to start), 4-bit codes, 2 per byte
  int tb_len;
};

// the ordered mode (no -M, the reference's default): the genome map built from the hit map,
// greedy (p.greedy, -G) or contextual, and its coordinate-ordered SAM (genome_map.c:186-800)
int write_sam_ordered(const char* path, const std::vector<HostRef>& refs, const HostReads& reads,
                      const std::vector<HostSlot>& slots, const storm_params& p, int argc, char** argv,
                      std::string& err);

// -O: the reads whose hit map row is empty (slots[4*r].ref_id == -1), as FASTQ in read order
// (hit_map__generate_unmapped_FASTQ_output, hit_map.c:514-578)
int write_unmapped_fastq(const char* path, const HostReads& reads, const std::vector<HostSlot>& slots,
                         const storm_params& p, std::string& err);

// SAM text of the ordered writer: header lines, MAPQ, and the fields of a mapped record (the -M
// records are formatted on the device, k_sam.hip)
struct SamFormat {
  SamFormat(const std::vector<HostRef>& refs, const HostReads& reads, const storm_params& p);
  void header(FILE* out, const char* sort_order, int argc, char** argv) const;
  int mapq(const HostSlot* row, int k) const;
  // QNAME .. QUAL and the AS / NM tags of rank k of read id, no line end
  void record(std::string& line, int64_t id, const HostSlot* row, int k, int flag, std::string& cigar,
              std::string& seq) const;
  const std::vector<HostRef>& refs;
  const HostReads& reads;
  const storm_params& p;
  bool color;
  std::vector<std::vector<uint8_t>> sym;   // color: base before each position of each reference
};

// the SAM header lines (hit_map.c:297-315): @HD with sort_order, @SQ per sequence, @RG, @PG with argv
void sam_header(FILE* out, const std::vector<HostRef>& refs, bool color, const std::string& db_name,
                const char* sort_order, int argc, char** argv);
// MAPQ of a record whose score is `diff` above its comparison score (hit_map.c:374-399)
int sam_mapq(int diff, int delta_m);

// items [0, n) formatted by fmt(item, text, cigar scratch, seq scratch) on up to 16 threads in
// blocks, written to out in item order; false when a write failed
bool write_blocks(FILE* out, int64_t n, const std::function<void(int64_t, std::string&, std::string&, std::string&)>& fmt);

uint32_t glibc_rand_next(uint32_t* state /* 35+ words, state[34] = position */);

}  // namespace storm
