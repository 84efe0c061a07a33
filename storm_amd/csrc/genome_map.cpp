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
#include <thread>
#include <unordered_map>

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

// Consensus counters of one reference sequence, positions 0..size (an insertion can follow the
// last code).  A position "exists" once a read was placed at it or counted over it; counters of a
// position that does not exist are never read.  Per position the reference keeps one counter word
// for the aligned codes and one per insertion offset 0..I-1, and every non-insertion cell of a
// placed read adds a gap (NO_CODE) to all I insertion counters (genome_map.c:512-515).  Each 6-bit
// counter is independent modulo 64, so that full-range gap count is kept once per position
// (`gaps`), and the few positions an insertion touches keep their per-offset counts in a side
// table; an insertion counter's gap count is the sum of the two, modulo 64.  The aligned counters
// use bits 0-29 of a word whose bit 31 marks the position as existing.
struct Consensus {
  static constexpr int P = 1 << PAGE_BITS;
  static constexpr uint32_t EXISTS = 0x80000000u;
  struct Page {
    uint32_t word[P];
    uint8_t gaps[P];
  };
  struct Ins {
    uint8_t code[32][4];   // per insertion offset, counts of the inserted codes 0..3 (mod 64)
    uint8_t gap[32];       // per insertion offset, gap counts beyond `gaps` (mod 64)
  };
  int64_t size = 0;
  int I = 0;
  std::vector<std::unique_ptr<Page>> pages;
  std::unordered_map<int64_t, Ins> ins;

  Consensus(int64_t size_, int I_) : size(size_), I(I_), pages((size_t)((size_ + 1 + P - 1) >> PAGE_BITS)) {}

  bool in_range(int64_t pos) const { return pos >= 0 && pos <= size; }
  const Page* page(int64_t pos) const { return pages[(size_t)(pos >> PAGE_BITS)].get(); }
  // the page holding pos, allocated on first use (positions in it do not exist until touched)
  Page* page_w(int64_t pos) {
    auto& pg = pages[(size_t)(pos >> PAGE_BITS)];
    if (!pg) pg.reset(new Page());   // value-initialised: all zero
    return pg.get();
  }
  // creates the position (genome_map__create_position_data, genome_map.c:68-99)
  Page* touch(int64_t pos) {
    auto& pg = pages[(size_t)(pos >> PAGE_BITS)];
    if (!pg) pg.reset(new Page());   // value-initialised: all zero
    pg->word[pos & (P - 1)] |= EXISTS;
    return pg.get();
  }
  // Whether the consensus code at pos (offset -1: aligned, else insertion offset), asked with x as
  // the default code, is x (genome_map__get_direct_consensus_code, genome_map.c:308-330, as its
  // callers use it, genome_map.c:373-420: every call passes the code it then compares with).  The
  // answer is x when the position does not exist; otherwise the code with the largest count, x on
  // a tie, and NO_CODE when every count is zero: x exactly when x's count is the largest and not
  // zero, or every count is zero and x is NO_CODE.  pg = page(pos).
  bool agrees(const Page* pg, int64_t pos, int offset, int x) const {
    const size_t o = (size_t)(pos & (P - 1));
    if (!pg || !(pg->word[o] & EXISTS)) return true;
    unsigned n[NO_CODE + 1];
    if (offset < 0) {
      const uint32_t w = pg->word[o];
      for (int c = 0; c <= NO_CODE; ++c) n[c] = cnt(w, c);
    } else {
      auto it = ins.find(pos);
      for (int c = 0; c < NO_CODE; ++c) n[c] = it == ins.end() ? 0u : it->second.code[offset][c];
      n[NO_CODE] = (pg->gaps[o] + (it == ins.end() ? 0u : it->second.gap[offset])) & 63u;
    }
    unsigned m = n[0];
    for (int c = 1; c <= NO_CODE; ++c) m = n[c] > m ? n[c] : m;
    return m == 0 ? x == NO_CODE : n[x] == m;
  }
  // genome_map__update_consensus_code (genome_map.c:340-349) for an inserted code at offset q
  void count_inserted(int64_t pos, int q, int code) {
    touch(pos);
    uint8_t& c = ins[pos].code[q][code];
    c = (uint8_t)((c + 1) & 63);
  }
  // the gap counts a cell adds at offsets q0..I-1 (genome_map.c:512-515)
  void count_gaps(int64_t pos, int q0) {
    if (q0 >= I) return;
    Page* pg = touch(pos);
    if (q0 == 0) {
      uint8_t& g = pg->gaps[pos & (P - 1)];
      g = (uint8_t)((g + 1) & 63);
      return;
    }
    Ins& e = ins[pos];
    for (int q = q0; q < I; ++q) e.gap[q] = (uint8_t)((e.gap[q] + 1) & 63);
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
  if (I < 0 || I > 31) { err = "-i outside 0..31"; return STORM_ERR_INVALID; }
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
    std::vector<uint64_t> keyed;
    for (int64_t id = 0; id < reads.n; ++id)
      if (slots[(size_t)(4 * id)].ref_id != -1)
        keyed.push_back(((uint64_t)(0xffffffffu - (uint32_t)slots[(size_t)(4 * id)].score) << 32) | (uint64_t)id);
    std::sort(keyed.begin(), keyed.end());   // keys are unique: this is the stable order
    // adjusted score of each hit of read id against the current consensus (genome_map.c:359-425);
    // false (bad = the position) when a hit leaves the sequence
    auto score_read = [&](int32_t id, int* adj, int& best_rank, int64_t& bad) -> bool {
      const HostSlot* row = &slots[(size_t)(4 * id)];
      for (int k = 0; k < 4; ++k) adj[k] = kMinScore;
      int best = kMinScore;
      best_rank = -1;
      for (int k = 0; k < 4 && row[k].ref_start != -1; ++k) {
        const HostSlot& h = row[k];
        const Consensus& cs = cons[(size_t)h.ref_id];
        const HostRef& ref = refs[(size_t)h.ref_id];
        int score = 0, indel = 0;
        int64_t pos = h.ref_start, pg_id = -1;
        const Consensus::Page* pg = nullptr;
        for (int c = 0; c < h.tb_len; ++c) {
          const int code = tb_code(h, c);
          const int t = pair_type(code, color);
          if (t < 0) continue;
          if (!cs.in_range(pos)) { bad = pos; return false; }
          if ((pos >> PAGE_BITS) != pg_id) { pg_id = pos >> PAGE_BITS; pg = cs.page(pos); }
          if (t == 0xC) {
            score += cs.agrees(pg, pos, indel, code & 3) ? p.match : p.mismatch;
            ++indel;
            continue;
          }
          const int x = t == 0x3 ? ref_code(ref, pos) : t == 0x1 ? NO_CODE : (code & 3);
          score += cs.agrees(pg, pos, -1, x) ? p.match : p.mismatch;
          ++pos;
          indel = 0;
        }
        adj[k] = score;
        if (score > best) { best = score; best_rank = k; }
      }
      return true;
    };
    // the positions [lo, hi] a hit's traceback reads or counts over: from its start, one position
    // per aligned or deleted code, and the one after them (insertions sit at the next position)
    auto span = [&](const HostSlot& h, int64_t& lo, int64_t& hi) {
      int64_t n = 0;
      for (int c = 0; c < h.tb_len; ++c) {
        const int t = pair_type(tb_code(h, c), color);
        n += (t >= 0 && t != 0xC);
      }
      lo = h.ref_start;
      hi = h.ref_start + n;
    };
    // Placement in chunks of reads: the adjusted scores of a chunk are computed in parallel
    // against the consensus at the chunk's start, then the reads are placed in order; a read any
    // of whose hits overlaps a position counted by an earlier read of the chunk is scored again
    // first.  A score that reads no changed counter is the one the sequential order computes.
    std::vector<std::vector<uint64_t>> dirty(refs.size());
    for (size_t r = 0; r < refs.size(); ++r) dirty[r].assign((size_t)((refs[r].size + 1 + 64) >> 6) + 1, 0);
    std::vector<std::pair<int32_t, std::pair<int64_t, int64_t>>> marked;
    auto is_dirty = [&](int32_t rid, int64_t lo, int64_t hi) {
      const std::vector<uint64_t>& d = dirty[(size_t)rid];
      for (int64_t w = lo >> 6; w <= (hi >> 6); ++w) {
        uint64_t m = d[(size_t)w];
        if (w == (lo >> 6)) m &= ~0ull << (lo & 63);
        if (w == (hi >> 6)) m &= (hi & 63) == 63 ? ~0ull : ((1ull << ((hi & 63) + 1)) - 1);
        if (m) return true;
      }
      return false;
    };
    auto mark = [&](int32_t rid, int64_t lo, int64_t hi) {
      std::vector<uint64_t>& d = dirty[(size_t)rid];
      for (int64_t w = lo >> 6; w <= (hi >> 6); ++w) {   // the bits lo..hi, a word at a time
        uint64_t m = ~0ull;
        if (w == (lo >> 6)) m &= ~0ull << (lo & 63);
        if (w == (hi >> 6)) m &= (hi & 63) == 63 ? ~0ull : ((1ull << ((hi & 63) + 1)) - 1);
        d[(size_t)w] |= m;
      }
      marked.push_back({rid, {lo, hi}});
    };
    const int64_t n_sorted = (int64_t)keyed.size();
    const int64_t B = 1 << 15;
    const int nt = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
    std::vector<int> sadj((size_t)(4 * B)), sbest((size_t)B);
    std::vector<uint8_t> sok((size_t)B);
    std::vector<int64_t> sbad((size_t)B), slo((size_t)(4 * B)), shi((size_t)(4 * B));
    for (int64_t s0 = 0; s0 < n_sorted; s0 += B) {
      const int64_t nb = std::min(B, n_sorted - s0);
      // per read of the chunk: its speculative scores, and each hit's span clipped to its sequence
      // (a span depends on the hit alone, so the ordered pass below only reads it)
      auto prep = [&](int64_t j) {
        const int32_t id = (int32_t)(keyed[(size_t)(s0 + j)] & 0xffffffffu);
        sok[(size_t)j] = score_read(id, &sadj[(size_t)(4 * j)], sbest[(size_t)j], sbad[(size_t)j]);
        const HostSlot* row = &slots[(size_t)(4 * id)];
        for (int k = 0; k < 4 && row[k].ref_start != -1; ++k) {
          int64_t lo, hi;
          span(row[k], lo, hi);
          slo[(size_t)(4 * j + k)] = std::max<int64_t>(lo, 0);
          shi[(size_t)(4 * j + k)] = std::min<int64_t>(hi, refs[(size_t)row[k].ref_id].size);
        }
      };
      auto work = [&](int t) {
        const int64_t per = (nb + nt - 1) / nt, j0 = t * per, j1 = std::min(nb, j0 + per);
        for (int64_t j = j0; j < j1; ++j) prep(j);
      };
      if (nb >= 4096 && nt > 1) {
        std::vector<std::thread> th;
        for (int t = 1; t < nt; ++t) th.emplace_back(work, t);
        work(0);
        for (auto& x : th) x.join();
      } else {
        for (int64_t j = 0; j < nb; ++j) prep(j);
      }
      constexpr int64_t PD = 8;   // prefetch distance in reads: slot rows at 2 PD, the data they name at PD
      for (int64_t j = 0; j < nb; ++j) {
        if (j + 2 * PD < nb) __builtin_prefetch(&slots[(size_t)(4 * (keyed[(size_t)(s0 + j + 2 * PD)] & 0xffffffffu))]);
        if (j + PD < nb) {
          const HostSlot* rp = &slots[(size_t)(4 * (keyed[(size_t)(s0 + j + PD)] & 0xffffffffu))];
          for (int k = 0; k < 4 && rp[k].ref_start != -1; ++k)   // the overlap check's bitmap words
            __builtin_prefetch(&dirty[(size_t)rp[k].ref_id][(size_t)(std::min<int64_t>(slo[(size_t)(4 * (j + PD) + k)],
                                                                                   refs[(size_t)rp[k].ref_id].size) >> 6)]);
        }
        if (j + PD < nb && sbest[(size_t)(j + PD)] >= 0) {
          const HostSlot& hp = slots[(size_t)(4 * (keyed[(size_t)(s0 + j + PD)] & 0xffffffffu)) + sbest[(size_t)(j + PD)]];
          __builtin_prefetch(hp.tb.data());
          const Consensus& cp = cons[(size_t)hp.ref_id];
          if (hp.ref_start >= 0 && hp.ref_start + hp.tb_len <= cp.size) {
            for (int64_t q = hp.ref_start; q < hp.ref_start + hp.tb_len; q += 16) {
              const Consensus::Page* pp = cp.page(q);
              if (!pp) continue;
              __builtin_prefetch(&pp->word[q & (Consensus::P - 1)], 1);
              if ((q - hp.ref_start) % 64 == 0) __builtin_prefetch(&pp->gaps[q & (Consensus::P - 1)], 1);
            }
            __builtin_prefetch(&refs[(size_t)hp.ref_id].code[(size_t)hp.ref_start]);
          }
        }
        const int32_t id = (int32_t)(keyed[(size_t)(s0 + j)] & 0xffffffffu);
        const HostSlot* row = &slots[(size_t)(4 * id)];
        int* adj = &sadj[(size_t)(4 * j)];
        int best_rank = sbest[(size_t)j];
        bool ok = sok[(size_t)j] != 0;
        int64_t bad = sbad[(size_t)j];
        bool redo = false;
        for (int k = 0; k < 4 && row[k].ref_start != -1 && !marked.empty() && !redo; ++k) {
          const int64_t lo = slo[(size_t)(4 * j + k)], hi = shi[(size_t)(4 * j + k)];
          redo = lo <= hi && is_dirty(row[k].ref_id, lo, hi);
        }
        if (redo) ok = score_read(id, adj, best_rank, bad);
        if (!ok) return out_of_range(id, bad);
        if (best_rank < 0 || !(p.uniqueness == 0 || adj[0] - adj[1] > p.uniqueness)) continue;
        // place the read and add its alignment to the consensus (genome_map.c:468-520)
        const HostSlot& h = row[best_rank];
        Consensus& cs = cons[(size_t)h.ref_id];
        const HostRef& ref = refs[(size_t)h.ref_id];
        int64_t pos = h.ref_start, pg_id = -1;
        if (!cs.in_range(pos)) return out_of_range(id, pos);
        cs.touch(pos);
        placed.push_back({h.ref_id, pos, id, best_rank});
        Consensus::Page* pg = nullptr;
        int ins = -1;
        for (int c = 0; c < h.tb_len; ++c) {
          const int code = tb_code(h, c);
          if (!cs.in_range(pos)) return out_of_range(id, pos);
          if ((pos >> PAGE_BITS) != pg_id) { pg_id = pos >> PAGE_BITS; pg = cs.page_w(pos); }
          uint32_t& w = pg->word[pos & (Consensus::P - 1)];
          switch (pair_type(code, color)) {   // the counts of genome_map.c:488-517, one cell
            case 0xC: ++ins; cs.count_inserted(pos, ins, code & 3); break;
            case 0x3: ins = -1; w |= Consensus::EXISTS; bump(w, ref_code(ref, pos)); break;
            case 0x4: case 0x8: ins = -1; w |= Consensus::EXISTS; bump(w, code & 3); break;
            case 0x1: ins = -1; w |= Consensus::EXISTS; bump(w, NO_CODE); break;
            default: break;
          }
          if (ins < 0 && I > 0) {   // a gap at every insertion offset
            w |= Consensus::EXISTS;
            uint8_t& gp = pg->gaps[pos & (Consensus::P - 1)];
            gp = (uint8_t)((gp + 1) & 63);
          } else {
            cs.count_gaps(pos, ins + 1);
          }
          if (ins < 0) ++pos;
        }
        mark(h.ref_id, h.ref_start, std::min<int64_t>(pos, cs.size));
      }
      for (auto& mr : marked) {
        std::vector<uint64_t>& d = dirty[(size_t)mr.first];
        for (int64_t w = mr.second.first >> 6; w <= (mr.second.second >> 6); ++w) d[(size_t)w] = 0;
      }
      marked.clear();
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
  const bool ok = write_blocks(out, (int64_t)placed.size(), [&](int64_t i, std::string& line, std::string& cigar, std::string& seq) {
    const Placement& pl = placed[(size_t)i];
    const HostSlot* row = &slots[(size_t)(4 * (int64_t)pl.read)];
    sf.record(line, pl.read, row, pl.rank, row[pl.rank].sense == -1 ? 0x10 : 0, cigar, seq);   // genome_map.c:688-776
    line += '\n';
  });
  if (fclose(out) != 0 || !ok) { err = std::string("error writing ") + path; return STORM_ERR_IO; }
  return STORM_OK;
}

}  // namespace storm
