// host_io.h -- host-side data of the mapping path: genome, reads, seeds, CSR index, SAM text.
#pragma once
#include <cstdint>
#include <cstdio>
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

// hit_map__generate_SAM_output (hit_map.c:197-504) for reads in order; slots[4*r + k]
int write_sam(const char* path, const std::vector<HostRef>& refs, const HostReads& reads,
              const std::vector<HostSlot>& slots, const storm_params& p, int argc, char** argv, std::string& err);

uint32_t glibc_rand_next(uint32_t* state /* 35+ words, state[34] = position */);

}  // namespace storm
