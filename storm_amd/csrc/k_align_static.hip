// k_align_static.hip -- the DP + traceback of records that a caller holds on the device, against a
// storm_align_plan's resident genome and reads, with the direction words in a static device array
// instead of memory the library allocates (storm_align_run_records_noalloc).
//
// For callers that cannot allocate device memory (code under HIP graph capture, tight serving
// loops): the call makes no hipMalloc, hipFree or other allocating call.  Every call on a GPU
// shares that GPU's array (each device holds its own copy, at its own address, until the process
// ends), so calls that overlap on one GPU race; calls on one stream, or waited for, do not.  The
// direction words are written before they are read within each record, so nothing needs clearing
// between calls.
#include <mutex>

#include "mapper.h"

namespace storm {

constexpr size_t ALIGN_STATIC_SCRATCH_BYTES = 16u << 20;
constexpr int MAX_DEVICES = 64;

}  // namespace storm

// the direction words of one launch: whole wavefronts of 64 records x (maxL x dp_dir_words(I) x 4)
// bytes, interleaved across the wavefront's lanes as k_align writes them
static __device__ __align__(256) uint8_t scratch[storm::ALIGN_STATIC_SCRATCH_BYTES];

namespace storm {
namespace {

std::mutex g_scratch_mu;
uint8_t* g_scratch_dev[MAX_DEVICES] = {};

// the device address of `scratch` on `device` (the current device), looked up once per device
int static_scratch(int device, uint8_t** out) {
  if (device < 0 || device >= MAX_DEVICES) return fail(STORM_ERR_INVALID, "device %d outside 0..%d", device, MAX_DEVICES - 1);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(STORM_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    g_scratch_dev[device] = (uint8_t*)p;
  }
  *out = g_scratch_dev[device];
  return STORM_OK;
}

}  // namespace
}  // namespace storm

using namespace storm;

extern "C" int storm_align_run_records_noalloc(storm_align_plan* plan, const int32_t* d_rec_rs, const int32_t* d_rec_diag,
                                               int64_t n_rec, storm_align_result* d_results, uint8_t* d_tb,
                                               void* stream) {
  AlignArgs a{};
  int maxL = 0, device = 0;
  hipStream_t own = nullptr;
  int rc = align_plan_view(plan, a, maxL, device, own);
  if (rc) return rc;
  if (n_rec < 0) return fail(STORM_ERR_INVALID, "negative record count");
  if (n_rec == 0) return STORM_OK;
  if (!d_rec_rs || !d_rec_diag || !d_results || !d_tb) return fail(STORM_ERR_INVALID, "null device buffer");
  if (((uintptr_t)d_tb & 3) || ((uintptr_t)d_results & 3))   // tracebacks are stored as 32-bit words
    return fail(STORM_ERR_INVALID, "d_tb and d_results must be 4-byte aligned");
  const int64_t stride = (int64_t)std::max(maxL, 1) * dp_dir_words(a.sc.I) * 4;
  const int64_t chunk = ((int64_t)ALIGN_STATIC_SCRATCH_BYTES / stride) & ~(int64_t)(ALIGN_WAVE_RECORDS - 1);
  if (chunk < ALIGN_WAVE_RECORDS)
    return fail(STORM_ERR_INVALID, "reads of %d codes at -i %d need %lld bytes of direction words per wavefront, "
                "more than the %zu-byte static scratch", maxL, a.sc.I, (long long)(ALIGN_WAVE_RECORDS * stride),
                ALIGN_STATIC_SCRATCH_BYTES);
  HIPOK(hipSetDevice(device));
  uint8_t* sp = nullptr;
  rc = static_scratch(device, &sp);
  if (rc) return rc;
  hipStream_t st = stream ? (hipStream_t)stream : own;
  a.rec_rs = d_rec_rs;
  a.rec_diag = d_rec_diag;
  a.out = (AlignRec*)d_results;
  a.tb = d_tb;
  a.scratch = sp;
  a.scratch_stride = stride;
  a.check_bounds = 1;
  // launches over the records in chunks that fit the array, one after another on one stream
  for (int64_t b0 = 0; b0 < n_rec; b0 += chunk) {
    a.base = b0;
    a.n = std::min(chunk, n_rec - b0);
    HIPOK(launch_align(a, st));
  }
  if (!stream) HIPOK(hipStreamSynchronize(st));
  return STORM_OK;
}
