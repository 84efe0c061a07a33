// genome_map.cpp -- the ordered mapping mode (no -M, the reference's default): each mapped read is
// placed at one of its hit map positions, greedily (-G: its best hit, genome_map.c:186-232) or in
// order of decreasing best score at the hit whose alignment agrees best with the consensus of the
// reads placed before it (genome_map.c:240-536); the SAM lists the placed reads by position
// (genome_map.c:551-800).  Host code: the placement is sequential by construction (every read
// sees the consensus left by all higher-scoring reads), and it runs once per run on the hit map
// the device produced.
#include <algorithm>
#include <cstring>
#include <memory>

#include "host_io.h"

namespace storm {

namespace {

constexpr int NO_CODE = 4;          // the gap / no-code symbol (GENOME_MAP_NO_CODE = CODE_COUNT)
constexpr int kMinScore = -16383;   // MIN_SCORE (alignment.h:47)
constexpr int PAGE_BITS = 12;       // positions per lazily allocated page of consensus counters

// 6-bit counters of codes 0..3 and of NO_CODE packed in one word, incremented modulo 64 like the
// reference's compressed counters (COUNTER_INC, genome_map.c:38-47)
inline unsigned cnt(uint32_t w, int code) { return (w >> (6 * code)) & 63u; }
inline void bump(uint32_t& w, int code) {
  const uint32_t u = (cnt(w, code) + 1) & 63u;
  w = (w & ~(63u << (6 * code))) | (u << (6 * code));
}

// consensus counters of one reference sequence, positions 0..size (an insertion can follow the
// last code): a position "exists" once a read was placed at it or counted over it; counters of a
// position that does not exist are never read
struct Consensus {
  struct Page {
    std::vector<uint8_t> exists;
    std::vector<uint32_t> aligned, inserted;   // inserted: I words per position
  };
  int64_t size = 0;
  int I = 0;
  std::vector<std::unique_ptr<Page>> pages;

  Consensus(int64_t size_, int I_) : size(size_), I(I_), pages((size_t)((size_ + 1 + (1 << PAGE_BITS) - 1) >> PAGE_BITS)) {}

  bool in_range(int64_t pos) const { return pos >= 0 && pos <= size; }
  Page* page(int64_t pos) const { return pages[(size_t)(pos >> PAGE_BITS)].get(); }
  // creates the position (genome_map__create_position_data, genome_map.c:68-99)
  void touch(int64_t pos) {
    auto& pg = pages[(size_t)(pos >> PAGE_BITS)];
    if (!pg) {
      pg.reset(new Page());
      pg->exists.assign((size_t)1 << PAGE_BITS, 0);
      pg->aligned.assign((size_t)1 << PAGE_BITS, 0);
      pg->inserted.assign(((size_t)1 << PAGE_BITS) * (size_t)I, 0);
    }
    pg->exists[(size_t)(pos & ((1 << PAGE_BITS) - 1))] = 1;
  }
  // the consensus code at pos (offset -1: aligned, else insertion offset) with reference_code as
  // the default and the tie winner (genome_map__get_direct_consensus_code, genome_map.c:308-330)
  int code(int64_t pos, int offset, int reference_code) const {
    const Page* pg = page(pos);
    const size_t o = (size_t)(pos & ((1 << PAGE_BITS) - 1));
    if (!pg || !pg->exists[o]) return reference_code;
    const uint32_t w = offset < 0 ? pg->aligned[o] : pg->inserted[o * (size_t)I + (size_t)offset];
    int best = reference_code;
    for (int c = 0; c <= NO_CODE; ++c)
      if (cnt(w, c) > cnt(w, best)) best = c;
    return cnt(w, best) == 0 ? NO_CODE : best;
  }
  // genome_map__update_consensus_code (genome_map.c:340-349)
  void count(int64_t pos, int offset, int code) {
    touch(pos);
    Page* pg = page(pos);
    const size_t o = (size_t)(pos & ((1 << PAGE_BITS) - 1));
    bump(offset < 0 ? pg->aligned[o] : pg->inserted[o * (size_t)I + (size_t)offset], code);
  }
};

// code k of a hit's traceback in alignment order (start first); the slot holds it end first
inline int tb_code(const HostSlot& h, int k) {
  const int n = h.tb_len - 1 - k;
  return (h.tb[(size_t)(n >> 1)] >> ((n & 1) ? 0 : 4)) & 0xF;
}

// GET_PAIR_TYPE (alignment.h:87-94): 0x4 code mismatch, 0x8 color-base mismatch (color only),
// 0xC insertion, else the code itself (0x3 match, 0x1 deletion); -1 for a code none of the
// reference's cases takes (genome_map.c:373-420, :488-517)
inline int pair_type(int code, bool color) {
  const int hi = code & 0xC;
  if (hi == 0x4) return 0x4;
  if (hi == 0x8) return color ? 0x8 : -1;
  if (hi == 0xC) return 0xC;
  return (code == 0x3 || code == 0x1) ? code : -1;
}

// the sequence's code at pos; the position after the last code reads the zero padding
inline int ref_code(const HostRef& ref, int64_t pos) { return pos < ref.size ? ref.code[(size_t)pos] : 0; }

struct Placement {
  int32_t ref_id;
  int64_t pos;
  int32_t read, rank;
};

}  // namespace

int write_sam_ordered(const char* path, const std::vector<HostRef>& refs, const HostReads& reads,
                      const std::vector<HostSlot>& slots, const storm_params& p, int argc, char** argv,
                      std::string& err) {
  const bool color = p.color != 0;
  const int I = p.allowed_indels;   // the hit map's indel_count (hit_map__create, run.c:818)
  std::vector<Placement> placed;
  auto out_of_range = [&](int64_t id, int64_t pos) {
    err = "read " + std::to_string(id) + ": genome map position " + std::to_string(pos) + " outside the sequence";
    return STORM_ERR_INVALID;
  };
  if (p.greedy) {   // genome_map__map_read_greedy (genome_map.c:186-201), reads in order
    for (int64_t id = 0; id < reads.n; ++id) {
      const HostSlot* row = &slots[(size_t)(4 * id)];
      if (row[0].ref_start == -1) continue;
      if (p.uniqueness == 0 || row[0].score - row[1].score >= p.uniqueness) {
        if (row[0].ref_start < 0 || row[0].ref_start > refs[(size_t)row[0].ref_id].size)
          return out_of_range(id, row[0].ref_start);
        placed.push_back({row[0].ref_id, row[0].ref_start, (int32_t)id, 0});
      }
    }
  } else {   // genome_map__build_contextual (genome_map.c:432-536)
    std::vector<Consensus> cons;
    cons.reserve(refs.size());
    for (auto& r : refs) cons.emplace_back(r.size, I);
    // mapped reads by decreasing best score, ties in read order: the reference's LSD radix sort
    // over hex digits with reversed buckets (genome_map.c:252-298) orders by the score's unsigned
    // 32-bit value
    std::vector<int32_t> order;
    for (int64_t id = 0; id < reads.n; ++id)
      if (slots[(size_t)(4 * id)].ref_id != -1) order.push_back((int32_t)id);
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
      return (uint32_t)slots[(size_t)(4 * a)].score > (uint32_t)slots[(size_t)(4 * b)].score;
    });
    for (const int32_t id : order) {
      const HostSlot* row = &slots[(size_t)(4 * id)];
      // adjusted score of each hit against the current consensus (genome_map.c:359-425)
      int adj[4] = {kMinScore, kMinScore, kMinScore, kMinScore};
      int best_rank = -1, best = kMinScore;
      for (int k = 0; k < 4 && row[k].ref_start != -1; ++k) {
        const HostSlot& h = row[k];
        const Consensus& cs = cons[(size_t)h.ref_id];
        const HostRef& ref = refs[(size_t)h.ref_id];
        int score = 0, indel = 0;
        int64_t pos = h.ref_start;
        for (int c = 0; c < h.tb_len; ++c) {
          const int code = tb_code(h, c);
          const int t = pair_type(code, color);
          if (t < 0) continue;
          if (!cs.in_range(pos)) return out_of_range(id, pos);
          if (t == 0xC) {
            score += ((code & 3) == cs.code(pos, indel, code & 3)) ? p.match : p.mismatch;
            ++indel;
            continue;
          }
          if (t == 0x3) {
            const int rc = ref_code(ref, pos);
            score += (rc == cs.code(pos, -1, rc)) ? p.match : p.mismatch;
          } else if (t == 0x1) {
            score += (cs.code(pos, -1, NO_CODE) == NO_CODE) ? p.match : p.mismatch;
          } else {
            score += ((code & 3) == cs.code(pos, -1, code & 3)) ? p.match : p.mismatch;
          }
          ++pos;
          indel = 0;
        }
        adj[k] = score;
        if (score > best) { best = score; best_rank = k; }
      }
      if (best_rank < 0 || !(p.uniqueness == 0 || adj[0] - adj[1] > p.uniqueness)) continue;
      // place the read and add its alignment to the consensus (genome_map.c:468-520)
      const HostSlot& h = row[best_rank];
      Consensus& cs = cons[(size_t)h.ref_id];
      const HostRef& ref = refs[(size_t)h.ref_id];
      int64_t pos = h.ref_start;
      if (!cs.in_range(pos)) return out_of_range(id, pos);
      cs.touch(pos);
      placed.push_back({h.ref_id, pos, id, best_rank});
      int ins = -1;
      for (int c = 0; c < h.tb_len; ++c) {
        const int code = tb_code(h, c);
        if (!cs.in_range(pos)) return out_of_range(id, pos);
        switch (pair_type(code, color)) {
          case 0xC: ++ins; cs.count(pos, ins, code & 3); break;
          case 0x3: ins = -1; cs.count(pos, -1, ref_code(ref, pos)); break;
          case 0x4: case 0x8: ins = -1; cs.count(pos, -1, code & 3); break;
          case 0x1: ins = -1; cs.count(pos, -1, NO_CODE); break;
          default: break;
        }
        for (int q = ins + 1; q < I; ++q) cs.count(pos, q, NO_CODE);
        if (ins < 0) ++pos;
      }
    }
  }
  // by reference, then position; reads at one position in placement order
  std::stable_sort(placed.begin(), placed.end(), [](const Placement& a, const Placement& b) {
    return a.ref_id != b.ref_id ? a.ref_id < b.ref_id : a.pos < b.pos;
  });
  FILE* out = fopen(path, "wb");
  if (!out) { err = std::string("cannot write ") + path; return STORM_ERR_IO; }
  const SamFormat sf(refs, reads, p);
  sf.header(out, "coordinate", argc, argv);
  write_blocks(out, (int64_t)placed.size(), [&](int64_t i, std::string& line, std::string& cigar, std::string& seq) {
    const Placement& pl = placed[(size_t)i];
    const HostSlot* row = &slots[(size_t)(4 * (int64_t)pl.read)];
    sf.record(line, pl.read, row, pl.rank, row[pl.rank].sense == -1 ? 0x10 : 0, cigar, seq);   // genome_map.c:688-776
    line += '\n';
  });
  if (fclose(out) != 0) { err = std::string("error writing ") + path; return STORM_ERR_IO; }
  return STORM_OK;
}

}  // namespace storm
