// storm_dev.h -- device-side data layout and helpers shared by the gfx950 kernels.
//
// HBM layout (all little-endian 32-bit words):
//   genome   : bit planes of 32 codes per word (lo = code bit 0, hi = code bit 1, mask = 1 when
//              masked), with GPAD codes of padding before code 0 and after the last code.  lo and
//              hi are interleaved ({lo, hi} as one 8-byte pair per 32 codes), so the window of a
//              filter candidate is one contiguous run of pairs; mask is a plane of its own (only the
//              DP reads it).  Padding codes are 0 (A) and masked, as in the reference buffers
//              (reference_data.c:75-88,184-188).
//   reads    : per read and strand, RW words of the lo plane then RW words of the hi plane
//              (row r, strand s at (2r+s)*2*RW), plus QW words of 2-bit quality levels.
//   index    : per seed, first[4^w + 1] and map[] (CSR, index.c:157-188), int32.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace storm {

constexpr int GPAD = 1024;          // padding codes on each side of a genome (multiple of 32)
constexpr int MAXL = 256;           // longest read handled by the kernels
constexpr int RW = MAXL / 32;       // words per read bit plane
constexpr int QW = MAXL / 16;       // words of 2-bit quality levels per read strand
constexpr int MIN_SCORE = -16383;   // alignment.h:47
constexpr int MAX_SEEDS = 8;
constexpr int MAX_SEED_WEIGHT = 14; // first[] holds 4^w + 1 entries
constexpr int MAX_I = 31;           // largest -i the DP kernel accepts

struct GenomeDev {
  const uint2* lh;          // {lo, hi} word pairs
  const uint32_t* mask;
  int64_t size;             // codes
};

struct ReadsDev {
  const uint32_t* planes;   // (2*n)*(2*RW) words
  const uint32_t* qual;     // (2*n)*QW words, 2-bit levels
  const int32_t* len;       // n
  int64_t n;
  int has_qual;
};

struct SeedDev {
  int weight, span, n_pos;
  int offs[32];             // offsets of the '#' symbols
  int pos[64];              // explicit read positions when n_pos > 0
  const int32_t* first;
  const int32_t* map;
};

struct SeedSetDev {
  int n;
  SeedDev s[MAX_SEEDS];
};

// scoring constants of the filter (alignment_simd.c:3640-3717) and the DP (alignment.c:255-285)
struct ScoreDev {
  int fmatch, fmismatch, fopen, fext;   // |m| |x| |o| |e| as u8
  int fthr;                             // min(z, 254)
  int ps[2][4];                         // pair_scores[match/mismatch][level]
  int gap_open, gap_ext;                // signed
  int color;
  int I;                                // -i
  int d;                                // min(I, 15)
  int W;                                // filter lanes
  int t;                                // -t
};

// 32 bits of the lo and hi planes starting at bit index `bit` (bit >= 0)
__device__ __forceinline__ void gbit_pair(const uint2* lh, int64_t bit, uint32_t& lo, uint32_t& hi) {
  const int64_t w = bit >> 5;
  const uint32_t s = (uint32_t)(bit & 31);
  const uint2 a = lh[w], b = lh[w + 1];
  lo = __builtin_amdgcn_alignbit(b.x, a.x, s);
  hi = __builtin_amdgcn_alignbit(b.y, a.y, s);
}

__device__ __forceinline__ int gcode(const GenomeDev& g, int64_t pos) {
  int64_t b = pos + GPAD;
  const uint2 p = g.lh[b >> 5];
  uint32_t l = (p.x >> (b & 31)) & 1u;
  uint32_t h = (p.y >> (b & 31)) & 1u;
  return (int)(l | (h << 1));
}

__device__ __forceinline__ int gmask(const GenomeDev& g, int64_t pos) {
  int64_t b = pos + GPAD;
  return (int)((g.mask[b >> 5] >> (b & 31)) & 1u);
}

__device__ __forceinline__ const uint32_t* read_planes(const ReadsDev& r, int64_t rs) {
  return r.planes + rs * (2 * RW);
}

__device__ __forceinline__ int rcode(const uint32_t* pl, int i) {
  return (int)(((pl[i >> 5] >> (i & 31)) & 1u) | (((pl[RW + (i >> 5)] >> (i & 31)) & 1u) << 1));
}

// The key of seed s at read position rp (seed.c:203-214): the codes at rp + offs[k], two bits
// each, first symbol highest.  A seed spanning at most 32 codes reads one 32-bit window per read
// plane (two words each, joined by alignbit) and takes its codes from registers; a longer one
// reads every code from the planes.
__device__ __forceinline__ int seed_key(const SeedDev& s, const uint32_t* pl, int rp) {
  int key = 0;
  const int w = rp >> 5, sh = rp & 31;
  if (s.span <= 32 && (w + 1 < RW || sh + s.span <= 32)) {
    // the second word: w + 1, or (last word of the plane, window inside it) w again, so no word
    // outside the read's planes is read
    const int w1 = w + 1 < RW ? w + 1 : w;
    const uint32_t lo = __builtin_amdgcn_alignbit(pl[w1], pl[w], sh);
    const uint32_t hi = __builtin_amdgcn_alignbit(pl[RW + w1], pl[RW + w], sh);
    for (int k = 0; k < s.weight; ++k) {
      const int o = s.offs[k];
      key = (key << 2) | (int)(((lo >> o) & 1u) | (((hi >> o) & 1u) << 1));
    }
    return key;
  }
  for (int k = 0; k < s.weight; ++k) key = (key << 2) | rcode(pl, rp + s.offs[k]);
  return key;
}

__device__ __forceinline__ int seed_npos(const SeedDev& s, int L) {
  if (s.n_pos > 0) return s.n_pos;
  const int n = L - s.span + 1;
  return n > 0 ? n : 0;
}

__device__ __forceinline__ int rqual(const ReadsDev& r, int64_t rs, int i) {
  if (!r.has_qual) return 3;   // alignment.c:151-162: no quality -> level of Q93
  return (int)((r.qual[rs * QW + (i >> 4)] >> ((i & 15) << 1)) & 3u);
}

}  // namespace storm
