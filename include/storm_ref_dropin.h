/*
 * storm_ref_dropin.h -- gfx950 entry points with the reference's own signatures and struct
 * layouts, for a build of the reference's C sources (compile with the reference's include path:
 * this header includes its index.h and alignment.h, and follows its NUCLEOTIDES switch).
 *
 * Library: storm_amd/_build/libstorm_hip.so (link with -lstorm_hip).  Every call runs on device 0.
 * The library mirrors the reference's ReferenceDBType, MaskedRegionType, SeedType, IndexType and
 * AlignmentType field by field (reference_data.h:11-53, seed.h:41-55, index.h:22-41,
 * alignment.h:131-229); storm_ref_abi_layout() reports the offsets it assumes.
 */
#ifndef STORM_REF_DROPIN_H
#define STORM_REF_DROPIN_H

#include <stdint.h>

#include "alignment.h"
#include "index.h"

#ifdef __cplusplus
extern "C" {
#endif

#ifdef NUCLEOTIDES
#define index_gfx950__build_reference index_gfx950__build_reference_nt
#define alignment_gfx950__align_batch alignment_gfx950__align_batch_nt
#else
#define index_gfx950__build_reference index_gfx950__build_reference_cs
#define alignment_gfx950__align_batch alignment_gfx950__align_batch_cs
#endif

/* Replaces index__build_reference (index.c:19-195, declared index.h:49; called per seed at
 * run.c:878-883): the index is built on the device (keys, masked-region skip of index.c:66-72,
 * stable sort by key, the repeat filter of index.c:117-156 with the program's FILTER_REPEATS /
 * ACCEPTED_STDEV_DISTANCE, index.c:9-10) and returned in the reference's contiguous layout
 * (index.c:157-188): first[index_size + 1] offsets, map[] positions ascending per key, both
 * malloc'd, so index__get_extern_next_hit (index.c:212-241) and index__destroy + free (run.c:989-991)
 * work on it unchanged.  Returns NULL on error (storm_last_error() has the message) where the
 * reference exits. */
IndexType *index_gfx950__build_reference(ReferenceDBType *db, SeedType *seed);

/* The same as index__destroy (index.c:255-260): frees first and map. */
void index_gfx950__destroy(IndexType *index);

/* alignment__align (alignment.c:488-632, declared alignment.h:280) followed, when do_traceback is
 * set and the score exceeds the threshold, by alignment__traceback (alignment.c:642-734,
 * alignment.h:289), for n alignments at once: each alignments[k] has been reset with its read and
 * window (alignment__reset_with_compressed_read / _ref, as process_read does, run.c:905,671) and
 * all share one parameter set (alignment__set_params, alignment.c:255-285).  scores[k] is
 * alignment__align(alignments[k], min_accepted[k]); best_i / best_j and the best cell's score in
 * the matrix (alignment__get_score(a, best_i, best_j), which hit_map__update reads), and with the traceback
 * best_i0 / best_j0 / traceback / traceback_seq_len, are set as those calls set them (the
 * displayed alignment, to_display, is not built).  Returns STORM_OK (0) or an error code. */
int alignment_gfx950__align_batch(AlignmentType *const *alignments, int n, const ScoreType *min_accepted,
                                  int do_traceback, ScoreType *scores);

/* offsets and sizes of the mirrored structs, in this order: MaskedRegionType {size, start, end,
 * next}, SeedType {size, weight, length, seed, positions, positions_count, mask}, IndexType {size,
 * db, seed, first, map, index_size, map_size}, AlignmentType {size, params, read_len, ref_len, read,
 * read_quality, reference, reference_masked, best_i, best_j, best_i0, best_j0, traceback,
 * traceback_seq_len, matrix, range_first}, AlignmentCellType {size, score}, AlignmentParamType
 * {size, gap, pair_scores, allowed_indels}, ReferenceDBType {size, name, size field, first_base
 * (-1 in the nucleotide build), sequence, sequence_masked, masked_regions}; color selects the storm-color ReferenceDBType.  Writes at
 * most n values and returns how many there are. */
int storm_ref_abi_layout(int color, int64_t *out, int n);

#ifdef __cplusplus
}
#endif
#endif /* STORM_REF_DROPIN_H */
