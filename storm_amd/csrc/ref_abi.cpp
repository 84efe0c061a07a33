// ref_abi.cpp -- entry points with the reference's own signatures and struct layouts
// (include/storm_ref_dropin.h), for a build of the reference's C sources that calls the GPU in
// place of its own index build and DP:
//   index_gfx950__build_reference  == index__build_reference (index.c:19-195, index.h:49)
//   index_gfx950__destroy          == index__destroy (index.c:255-260, index.h:80)
//   alignment_gfx950__align_batch  == alignment__align + alignment__traceback (alignment.c:488-734,
//                                     alignment.h:280,289) over many AlignmentType at once
// The reference's structs are mirrored field by field below (their layout is what the reference's
// code reads; storm_ref_abi_layout reports the offsets so a test can compare them with the
// reference's headers).  Two builds of the reference exist (storm-nucleotide with NUCLEOTIDES,
// storm-color without): ReferenceDBType differs between them and the DP follows the alphabet, so
// each entry point exists as _nt and _cs; the header picks one by NUCLEOTIDES.
#include <algorithm>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mapper.h"

using namespace storm;

namespace {

// reference_data.h:11-18
struct RefMaskedRegion {
  int start;
  int end;
  RefMaskedRegion* next;
};
// reference_data.h:25-53, storm-nucleotide
struct RefDBNt {
  char* name;
  int size;
  unsigned char* sequence_alloc;
  unsigned char* sequence;
  unsigned char* sequence_masked_alloc;
  unsigned char* sequence_masked;
  RefMaskedRegion* masked_regions;
};
// reference_data.h:25-53, storm-color (first_base after size)
struct RefDBCs {
  char* name;
  int size;
  unsigned char first_base;
  unsigned char* sequence_alloc;
  unsigned char* sequence;
  unsigned char* sequence_masked_alloc;
  unsigned char* sequence_masked;
  RefMaskedRegion* masked_regions;
};
// seed.h:41-55
struct RefSeed {
  int weight;
  int length;
  unsigned char* seed;
  int* positions;
  int positions_count;
  unsigned long long mask;
};
// index.h:22-41
struct RefIndex {
  void* db;
  RefSeed* seed;
  int* first;
  int* map;
  int index_size;
  int map_size;
};
// alignment.h:131-139 (GAP_EXTEND_IDX 0, GAP_OPEN_IDX 1; SYMBOL_PAIR_MATCH_IDX 0, _MISMATCH_IDX 1)
struct RefAlignParams {
  int gap[2];
  int pair_scores[2][4];
  short allowed_indels;
};
// alignment.h:189-229 (the same in both builds)
struct RefAlignment {
  RefAlignParams params;
  short read_len, ref_len;
  unsigned char* read;
  int* read_quality;
  unsigned char* reference;
  unsigned char* reference_masked;
  void* matrix_alloc;
  void* matrix;
  short* range_first;
  short range_shift;
  short best_i, best_j;
  short best_i0, best_j0;
  char* to_display;
  unsigned char* traceback;
  int traceback_seq_len;
};
// alignment.h:158-178: AlignmentCellType, storm-nucleotide and storm-color
struct RefCellNt {
  int score;
  char from;
  short indels;
};
struct RefCellCs {
  int score;
  unsigned char color_checksum;
  char from;
  short indels;
  short color_mismatch_freshness;
  unsigned char color_autocorrect;
};

}  // namespace

// The reference's repeat-filter settings (index.c:9-10, set from -l by main.c:413-418), read
// where the program defines them; a program without them gets the reference's defaults.
extern "C" int FILTER_REPEATS __attribute__((weak));
extern "C" int ACCEPTED_STDEV_DISTANCE __attribute__((weak));

namespace {

int repeat_distance() {
  const int on = &FILTER_REPEATS ? FILTER_REPEATS : 1;
  const int dist = &ACCEPTED_STDEV_DISTANCE ? ACCEPTED_STDEV_DISTANCE : 100;
  return on ? dist : 0;
}

template <class DB>
RefIndex* build_reference(DB* db, RefSeed* seed) {
  RefIndex* index = (RefIndex*)calloc(1, sizeof(RefIndex));
  if (!index || !db || !seed) {
    fail(STORM_ERR_INVALID, "index_gfx950__build_reference: null argument");
    free(index);
    return nullptr;
  }
  index->index_size = 1 << (seed->weight << 1);
  index->map_size = db->size - seed->length + 1;
  index->db = db;
  index->seed = seed;
  if (index->map_size < 0) {   // index.c:25-28: nothing indexed
    index->map_size = 0;
    return index;
  }
  // the seed pattern from its match mask (bit k = symbol k is a '#', seed.c:25-52)
  std::string pattern((size_t)seed->length, '-');
  for (int k = 0; k < seed->length && k < 64; ++k)
    if ((seed->mask >> k) & 1ull) pattern[(size_t)k] = '#';
  std::vector<int32_t> rs, re;
  for (const RefMaskedRegion* r = db->masked_regions; r; r = r->next) {
    rs.push_back(r->start);
    re.push_back(r->end);
  }
  storm_index* ix = nullptr;
  if (storm_index_build(0, db->sequence, db->size, rs.data(), re.data(), (int)rs.size(), pattern.c_str(),
                        repeat_distance(), &ix) != STORM_OK) {
    free(index);
    return nullptr;
  }
  int64_t n_keys = 0, n_map = 0;
  storm_index_sizes(ix, &n_keys, &n_map, nullptr, nullptr);
  // index.c:40-44,157-188: first[index_size + 1] offsets, map[map_size] with the grouped positions first
  index->first = (int*)malloc(sizeof(int) * (size_t)(index->index_size + 1));
  index->map = (int*)malloc(sizeof(int) * (size_t)std::max(index->map_size, 1));
  if (!index->first || !index->map || n_keys != index->index_size || n_map > index->map_size ||
      storm_index_fetch(ix, index->first, index->map) != STORM_OK) {
    if (!index->first || !index->map) fail(STORM_ERR_MEMORY, "index_gfx950__build_reference: out of memory");
    else if (n_keys != index->index_size || n_map > index->map_size)
      fail(STORM_ERR_DEVICE, "index_gfx950__build_reference: index sizes disagree");
    storm_index_destroy(ix);
    free(index->first);
    free(index->map);
    free(index);
    return nullptr;
  }
  storm_index_destroy(ix);
  return index;
}

void destroy(RefIndex* index) {   // index.c:255-260
  if (!index) return;
  free(index->first);
  index->first = nullptr;
  free(index->map);
  index->map = nullptr;
}

// The DP of n reset alignments (alignment__reset_with_compressed_read / _ref done by the caller, as
// process_read does, run.c:905,671): their windows laid end to end, with masked gaps, as one
// sequence on the device, their reads as read strands, one record per alignment; the kernel runs
// at the lowest threshold of the batch and reports the survival bound, from which each alignment's
// own early exit follows (alignment.c:613-616: the DP with threshold T stops exactly when the bound
// falls below T).
int align_batch(RefAlignment* const* als, int n, const int* min_accepted, int do_traceback, int* scores, int color) {
  if (n < 0 || (n > 0 && (!als || !min_accepted || !scores))) return fail(STORM_ERR_INVALID, "alignment_gfx950__align_batch: null argument");
  if (n == 0) return STORM_OK;
  const RefAlignParams& p0 = als[0]->params;
  const int I = p0.allowed_indels;
  for (int k = 0; k < n; ++k) {
    const RefAlignment* a = als[k];
    if (!a || memcmp(&a->params, &p0, offsetof(RefAlignParams, allowed_indels) + sizeof(short)) != 0)
      return fail(STORM_ERR_INVALID, "alignment %d: the batch needs one parameter set (scores, -i)", k);
    if (a->read_len < 1 || a->read_len > MAXL || a->ref_len != a->read_len + 2 * I)
      return fail(STORM_ERR_INVALID, "alignment %d: read length %d / window %d (1..%d, window = read + 2i)", k,
                  a->read_len, a->ref_len, MAXL);
  }
  // the scores must be the ones alignment__set_params derives (alignment.c:255-285)
  storm_params sp;
  storm_params_default(&sp, color);
  sp.match = p0.pair_scores[0][3];
  sp.mismatch = p0.pair_scores[1][3];
  sp.gap_open = p0.gap[1];
  sp.gap_extend = p0.gap[0];
  sp.allowed_indels = I;
  for (int q = 0; q < 4; ++q) {
    const int m = (int)((sp.match * (q + 1) + 0.) / 4), x = (int)((sp.mismatch * (q + 1) + 0.) / 4);   // SCALE_SCORE
    if (p0.pair_scores[0][q] != m || p0.pair_scores[1][q] != x)
      return fail(STORM_ERR_INVALID, "pair scores are not the quality scaling of match %d / mismatch %d", sp.match,
                  sp.mismatch);
  }
  int tmin = min_accepted[0];
  int maxL = 0;
  for (int k = 0; k < n; ++k) {
    tmin = std::min(tmin, min_accepted[k]);
    maxL = std::max<int>(maxL, als[k]->read_len);
  }
  sp.t = tmin;
  // windows laid out with 64 masked codes between them
  const int64_t S = (int64_t)maxL + 2 * I + 64;
  const int64_t G = 64 + (int64_t)n * S + 64;
  std::vector<uint8_t> gcodes((size_t)G, 0), gmask((size_t)G, 1);
  const int64_t stride = (maxL + 3) / 4;
  std::vector<uint8_t> rp((size_t)(n * stride) + 16, 0), qp((size_t)(n * stride) + 16, 0);
  std::vector<int32_t> rlen((size_t)n), rec_rs((size_t)n), rec_diag((size_t)n);
  for (int k = 0; k < n; ++k) {
    const RefAlignment* a = als[k];
    const int64_t o = 64 + (int64_t)k * S;
    for (int j = 0; j < a->ref_len; ++j) {
      gcodes[(size_t)(o + j)] = a->reference[j] & 3;
      gmask[(size_t)(o + j)] = a->reference_masked[j] ? 1 : 0;
    }
    uint8_t* r = &rp[(size_t)(k * stride)];
    uint8_t* q = &qp[(size_t)(k * stride)];
    for (int i = 0; i < a->read_len; ++i) {
      r[i >> 2] |= (uint8_t)((a->read[i] & 3) << ((i & 3) << 1));
      q[i >> 2] |= (uint8_t)((a->read_quality[i] & 3) << ((i & 3) << 1));
    }
    rlen[(size_t)k] = a->read_len;
    rec_rs[(size_t)k] = 2 * k;                  // the read as given (strand 0)
    rec_diag[(size_t)k] = (int32_t)(o + I);     // window start max(h - i, -min(i, 15)) = o
  }
  std::vector<uint8_t> gpacked((size_t)(G + 3) / 4, 0);
  for (int64_t j = 0; j < G; ++j) gpacked[(size_t)(j >> 2)] |= (uint8_t)(gcodes[(size_t)j] << ((j & 3) << 1));
  storm_align_plan* plan = nullptr;
  int rc = storm_align_plan_create(0, gpacked.data(), gmask.data(), G, rp.data(), stride, qp.data(), stride, nullptr,
                                   rlen.data(), n, rec_rs.data(), rec_diag.data(), n, &sp, &plan);
  if (rc) return rc;
  std::vector<storm_align_result> res((size_t)n);
  std::vector<uint8_t> tb((size_t)n * STORM_TB_BYTES);
  rc = storm_align_run(plan, 1, nullptr);
  if (!rc) rc = storm_align_fetch(plan, res.data(), tb.data());
  storm_align_plan_destroy(plan);
  if (rc) return rc;
  for (int k = 0; k < n; ++k) {
    RefAlignment* a = als[k];
    const storm_align_result& x = res[(size_t)k];
    const int T = min_accepted[k];
    if (x.score == MIN_SCORE || x.bound < T) { scores[k] = MIN_SCORE; continue; }   // the DP at T stopped
    scores[k] = x.score;
    if (x.best_j >= 0) {   // alignment.c:621-630
      a->best_i = (short)(a->read_len - 1);
      a->best_j = x.best_j;
      // the best cell's score, which alignment__get_score (alignment.c:90-92) reads back, as
      // hit_map__update does (hit_map.c:77)
      char* row = ((char**)a->matrix)[a->best_i] +
                  (ptrdiff_t)(a->best_j - a->range_first[a->best_i]) * (color ? sizeof(RefCellCs) : sizeof(RefCellNt));
      memcpy(row, &x.score, sizeof(int));
    }
    if (do_traceback && x.score > T) {   // alignment__traceback (alignment.c:642-734), as run.c:675-677
      a->best_i0 = x.best_i0;
      a->best_j0 = x.best_j0;
      a->traceback_seq_len = x.tb_len;
      memcpy(a->traceback, &tb[(size_t)k * STORM_TB_BYTES], (size_t)(x.tb_len + 1) / 2);
    }
  }
  return STORM_OK;
}

}  // namespace

extern "C" {

void* index_gfx950__build_reference_nt(void* db, void* seed) { return build_reference((RefDBNt*)db, (RefSeed*)seed); }
void* index_gfx950__build_reference_cs(void* db, void* seed) { return build_reference((RefDBCs*)db, (RefSeed*)seed); }
void index_gfx950__destroy(void* index) { destroy((RefIndex*)index); }

int alignment_gfx950__align_batch_nt(void* const* alignments, int n, const int* min_accepted, int do_traceback,
                                     int* scores) {
  return align_batch((RefAlignment* const*)alignments, n, min_accepted, do_traceback, scores, 0);
}
int alignment_gfx950__align_batch_cs(void* const* alignments, int n, const int* min_accepted, int do_traceback,
                                     int* scores) {
  return align_batch((RefAlignment* const*)alignments, n, min_accepted, do_traceback, scores, 1);
}

// offsets and sizes of the mirrored structs, for comparison with the reference's headers:
// out[0..n) in the order documented in storm_ref_dropin.h; returns the count written
int storm_ref_abi_layout(int color, int64_t* out, int n) {
  std::vector<int64_t> v = {
      (int64_t)sizeof(RefMaskedRegion), offsetof(RefMaskedRegion, start), offsetof(RefMaskedRegion, end),
      offsetof(RefMaskedRegion, next),
      (int64_t)sizeof(RefSeed), offsetof(RefSeed, weight), offsetof(RefSeed, length), offsetof(RefSeed, seed),
      offsetof(RefSeed, positions), offsetof(RefSeed, positions_count), offsetof(RefSeed, mask),
      (int64_t)sizeof(RefIndex), offsetof(RefIndex, db), offsetof(RefIndex, seed), offsetof(RefIndex, first),
      offsetof(RefIndex, map), offsetof(RefIndex, index_size), offsetof(RefIndex, map_size),
      (int64_t)sizeof(RefAlignment), offsetof(RefAlignment, params), offsetof(RefAlignment, read_len),
      offsetof(RefAlignment, ref_len), offsetof(RefAlignment, read), offsetof(RefAlignment, read_quality),
      offsetof(RefAlignment, reference), offsetof(RefAlignment, reference_masked), offsetof(RefAlignment, best_i),
      offsetof(RefAlignment, best_j), offsetof(RefAlignment, best_i0), offsetof(RefAlignment, best_j0),
      offsetof(RefAlignment, traceback), offsetof(RefAlignment, traceback_seq_len), offsetof(RefAlignment, matrix),
      offsetof(RefAlignment, range_first),
      (int64_t)(color ? sizeof(RefCellCs) : sizeof(RefCellNt)),
      color ? (int64_t)offsetof(RefCellCs, score) : (int64_t)offsetof(RefCellNt, score),
      (int64_t)sizeof(RefAlignParams), offsetof(RefAlignParams, gap), offsetof(RefAlignParams, pair_scores),
      offsetof(RefAlignParams, allowed_indels)};
  if (color) {
    const int64_t d[] = {(int64_t)sizeof(RefDBCs), offsetof(RefDBCs, name), offsetof(RefDBCs, size),
                         offsetof(RefDBCs, first_base), offsetof(RefDBCs, sequence), offsetof(RefDBCs, sequence_masked),
                         offsetof(RefDBCs, masked_regions)};
    v.insert(v.end(), std::begin(d), std::end(d));
  } else {
    const int64_t d[] = {(int64_t)sizeof(RefDBNt), offsetof(RefDBNt, name), offsetof(RefDBNt, size), -1,
                         offsetof(RefDBNt, sequence), offsetof(RefDBNt, sequence_masked),
                         offsetof(RefDBNt, masked_regions)};
    v.insert(v.end(), std::begin(d), std::end(d));
  }
  const int m = std::min<int>(n, (int)v.size());
  for (int k = 0; k < m; ++k) out[k] = v[(size_t)k];
  return (int)v.size();
}

}  // extern "C"
