// align_batch.cpp -- the banded DP and traceback of a batch of (read strand, diagonal) records
// with genome, reads and records resident in HBM (storm_align_*): alignment__align and
// alignment__traceback (alignment.c:326-734) as process_read calls them for each filter pass
// (run.c:665-690), run by the staged k_align kernel the plain mapping path uses.
#include <algorithm>
#include <cstring>

#include "mapper.h"

using namespace storm;

static_assert(sizeof(storm_align_result) == sizeof(AlignRec), "storm_align_result mirrors AlignRec");
static_assert(offsetof(storm_align_result, score) == offsetof(AlignRec, score), "layout");
static_assert(offsetof(storm_align_result, bound) == offsetof(AlignRec, S), "layout");
static_assert(offsetof(storm_align_result, win) == offsetof(AlignRec, win), "layout");
static_assert(offsetof(storm_align_result, best_j) == offsetof(AlignRec, best_j), "layout");
static_assert(offsetof(storm_align_result, tb_len) == offsetof(AlignRec, tb_len), "layout");
static_assert(STORM_TB_BYTES == TBB, "traceback bytes per record");

struct storm_align_plan {
  int device = 0;
  hipStream_t st = nullptr;
  DevGenome g;
  DevReads reads;
  DevMem rec_rs, rec_diag, recs, tb, scratch;
  AlignArgs a{};
  int maxL = 0;
  int64_t n = 0;
  ~storm_align_plan() {
    if (!st) return;
    (void)hipSetDevice(device);
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
  }
};

extern "C" {

int storm_align_plan_create(int device, const uint8_t* genome_packed, const uint8_t* genome_mask, int64_t genome_codes,
                            const uint8_t* reads_packed, int64_t read_stride, const uint8_t* qual, int64_t qual_stride,
                            const uint8_t* first_qual, const int32_t* read_len, int64_t n_reads, const int32_t* rec_rs,
                            const int32_t* rec_diag, int64_t n_rec, const storm_params* params,
                            storm_align_plan** out) {
  if (!out) return fail(STORM_ERR_INVALID, "no output pointer");
  *out = nullptr;
  if (!params) return fail(STORM_ERR_INVALID, "no parameters");
  const int I = params->allowed_indels;
  if (I < 0 || I > MAX_I) return fail(STORM_ERR_INVALID, "-i %d outside 0..%d", I, MAX_I);
  if (params->match <= 0 || params->mismatch > 0 || params->gap_open > 0 || params->gap_extend > 0)
    return fail(STORM_ERR_INVALID, "scores: match > 0, mismatch/gap penalties <= 0 (main.c:262-291)");
  if (genome_codes < 1 || !genome_packed) return fail(STORM_ERR_INVALID, "empty genome");
  if (n_reads < 0 || n_rec < 0) return fail(STORM_ERR_INVALID, "negative count");
  if (n_reads > 0 && (!reads_packed || !read_len)) return fail(STORM_ERR_INVALID, "no reads");
  if (n_rec > 0 && (!rec_rs || !rec_diag)) return fail(STORM_ERR_INVALID, "no records");
  if (n_reads > INT32_MAX / 2) return fail(STORM_ERR_INVALID, "too many reads");
  int maxL = 0;
  for (int64_t r = 0; r < n_reads; ++r) {
    const int L = read_len[r];
    if (L < 1 || L > MAXL)
      return fail(STORM_ERR_INVALID, "read %lld has length %d (1..%d supported)", (long long)r, L, MAXL);
    if ((int64_t)(L + 3) / 4 > read_stride) return fail(STORM_ERR_INVALID, "read_stride below a read's packed bytes");
    if (qual && (int64_t)(L + 3) / 4 > qual_stride) return fail(STORM_ERR_INVALID, "qual_stride below a read's packed levels");
    maxL = std::max(maxL, L);
  }
  const int d = I >= 16 ? 15 : I;
  for (int64_t k = 0; k < n_rec; ++k) {
    if (rec_rs[k] < 0 || rec_rs[k] >= 2 * n_reads)
      return fail(STORM_ERR_INVALID, "record %lld: read strand %d outside 0..%lld", (long long)k, rec_rs[k],
                  (long long)(2 * n_reads - 1));
    // window codes win .. win + L + 2I - 1, and the DP's look-ahead words, inside the padded planes
    const int64_t h = rec_diag[k];
    const int64_t win = (h - I > -d) ? h - I : -d;
    if (win + read_len[rec_rs[k] >> 1] + 2 * I + 96 > genome_codes + GPAD)
      return fail(STORM_ERR_INVALID, "record %lld: diagonal %lld runs past the genome end", (long long)k, (long long)h);
  }
  std::unique_ptr<storm_align_plan> pl(new storm_align_plan());
  pl->device = device;
  HIPOK(hipSetDevice(device));
  HIPOK(hipStreamCreateWithFlags(&pl->st, hipStreamNonBlocking));
  hipStream_t st = pl->st;
  std::vector<uint8_t> codes((size_t)genome_codes), mask;
  for (int64_t i = 0; i < genome_codes; ++i) codes[(size_t)i] = (genome_packed[i >> 2] >> ((i & 3) << 1)) & 3;
  if (genome_mask) {
    mask.resize((size_t)genome_codes);
    for (int64_t i = 0; i < genome_codes; ++i) mask[(size_t)i] = genome_mask[i] ? 1 : 0;
  }
  int rc = upload_genome(codes.data(), genome_mask ? mask.data() : nullptr, genome_codes, pl->g, st);
  if (rc) return rc;
  DevReads& dr = pl->reads;
  const int64_t nr = std::max<int64_t>(n_reads, 1);
  HIPOK(dr.codes.ensure((size_t)(nr * read_stride) + 16));
  HIPOK(dr.len.ensure((size_t)nr * 4));
  if (n_reads) {
    HIPOK(hipMemcpyAsync(dr.codes.p, reads_packed, (size_t)(n_reads * read_stride), hipMemcpyHostToDevice, st));
    HIPOK(hipMemcpyAsync(dr.len.p, read_len, (size_t)n_reads * 4, hipMemcpyHostToDevice, st));
  }
  if (qual) {
    HIPOK(dr.qual.ensure((size_t)(nr * qual_stride) + 16));
    if (n_reads) HIPOK(hipMemcpyAsync(dr.qual.p, qual, (size_t)(n_reads * qual_stride), hipMemcpyHostToDevice, st));
  }
  if (first_qual) {
    HIPOK(dr.first_qual.ensure((size_t)nr));
    if (n_reads) HIPOK(hipMemcpyAsync(dr.first_qual.p, first_qual, (size_t)n_reads, hipMemcpyHostToDevice, st));
  }
  dr.n = n_reads;
  dr.has_qual = qual ? 1 : 0;
  HIPOK(dr.planes.ensure((size_t)(2 * nr) * 2 * RW * 4));
  HIPOK(dr.qplanes.ensure((size_t)(2 * nr) * QW * 4));
  PrepArgs pa{};
  pa.codes = dr.codes.as<uint8_t>();
  pa.code_stride = (int)read_stride;
  pa.qual = qual ? dr.qual.as<uint8_t>() : nullptr;
  pa.qual_stride = (int)qual_stride;
  pa.first_qual = first_qual ? dr.first_qual.as<uint8_t>() : nullptr;
  pa.len = dr.len.as<int32_t>();
  pa.n = n_reads;
  pa.color = params->color ? 1 : 0;
  pa.planes = dr.planes.as<uint32_t>();
  pa.qplanes = dr.qplanes.as<uint32_t>();
  if (n_reads) HIPOK(launch_prep_reads(pa, st));
  const int64_t nc = std::max<int64_t>(n_rec, 1);
  HIPOK(pl->rec_rs.ensure((size_t)nc * 4));
  HIPOK(pl->rec_diag.ensure((size_t)nc * 4));
  HIPOK(pl->recs.ensure((size_t)nc * sizeof(AlignRec)));
  HIPOK(pl->tb.ensure((size_t)nc * TBB));
  if (n_rec) {
    HIPOK(hipMemcpyAsync(pl->rec_rs.p, rec_rs, (size_t)n_rec * 4, hipMemcpyHostToDevice, st));
    HIPOK(hipMemcpyAsync(pl->rec_diag.p, rec_diag, (size_t)n_rec * 4, hipMemcpyHostToDevice, st));
  }
  HIPOK(hipStreamSynchronize(st));   // the host staging vectors are released on return
  storm_params p = *params;
  p.color = p.color ? 1 : 0;
  AlignArgs& a = pl->a;
  fill_score(p, a.sc);
  a.g = pl->g.view();
  a.r = dr.view();
  a.rec_rs = pl->rec_rs.as<int32_t>();
  a.rec_diag = pl->rec_diag.as<int32_t>();
  a.out = pl->recs.as<AlignRec>();
  a.tb = pl->tb.as<uint8_t>();
  pl->maxL = maxL;
  pl->n = n_rec;
  // size the direction scratch now, so that storm_align_run allocates nothing
  HIPOK(pl->scratch.ensure((size_t)align_scratch_bytes(I, maxL, n_rec)));
  *out = pl.release();
  return STORM_OK;
}

int storm_align_run(storm_align_plan* pl, int reps, double* ms_per_run) {
  if (!pl || reps < 1) return fail(STORM_ERR_INVALID, "bad plan or reps");
  HIPOK(hipSetDevice(pl->device));
  hipEvent_t a, b;
  HIPOK(hipEventCreate(&a));
  HIPOK(hipEventCreate(&b));
  struct EvGuard { hipEvent_t x, y; ~EvGuard() { (void)hipEventDestroy(x); (void)hipEventDestroy(y); } } eg{a, b};
  HIPOK(hipEventRecord(a, pl->st));
  for (int k = 0; k < reps; ++k) {
    const int rc = align_records(pl->a, pl->maxL, pl->n, pl->scratch, pl->st);
    if (rc) return rc;
  }
  HIPOK(hipEventRecord(b, pl->st));
  HIPOK(hipEventSynchronize(b));
  float ms = 0;
  HIPOK(hipEventElapsedTime(&ms, a, b));
  if (ms_per_run) *ms_per_run = ms / reps;
  return STORM_OK;
}

int storm_align_fetch(storm_align_plan* pl, storm_align_result* results, uint8_t* tb) {
  if (!pl) return fail(STORM_ERR_INVALID, "bad plan");
  HIPOK(hipSetDevice(pl->device));
  if (results && pl->n)
    HIPOK(hipMemcpyAsync(results, pl->recs.p, (size_t)pl->n * sizeof(AlignRec), hipMemcpyDeviceToHost, pl->st));
  if (tb && pl->n) HIPOK(hipMemcpyAsync(tb, pl->tb.p, (size_t)pl->n * TBB, hipMemcpyDeviceToHost, pl->st));
  HIPOK(hipStreamSynchronize(pl->st));
  return STORM_OK;
}

void storm_align_plan_destroy(storm_align_plan* pl) { delete pl; }

}  // extern "C"

namespace storm {
int align_plan_view(storm_align_plan* pl, AlignArgs& a, int& maxL, int& device, hipStream_t& st) {
  if (!pl) return fail(STORM_ERR_INVALID, "bad plan");
  a = pl->a;
  maxL = pl->maxL;
  device = pl->device;
  st = pl->st;
  return STORM_OK;
}
}  // namespace storm
