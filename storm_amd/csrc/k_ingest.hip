// k_ingest.hip -- FASTQ reads parsed and packed on the device (load_reads_db, reads_data.c:351-566;
// string_to_compressed_code, codes.c:170-197; compress_quality_sequence, read_quality.c:18-41).
//
// The file's bytes are copied to HBM once.  Newlines are counted per 256-byte chunk, a prefix sum
// numbers them, and a second pass writes the byte offset of every newline, which gives every line
// its [start, end) range (trailing '\r' removed, as the host loader's line split does).  A FASTQ
// record is four lines; one thread per read checks its record, then packs its codes (4 per byte,
// code k at bits 2(k & 3) of byte k >> 2), its quality levels and its tag (the largest value of
// the last-32-codes window, reads_data.c:494-502).  The reads are then ordered by tag with a stable
// radix sort (the reference sorts them with qsort on the tag, reads_data.c:59-67).
#include <hip/hip_runtime.h>
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "storm_kernels.h"

namespace storm {

constexpr int64_t NL_CHUNK = 256;

__global__ void __launch_bounds__(256) k_nl_count(const uint8_t* __restrict__ d, int64_t n, int32_t* counts,
                                                  int64_t n_chunks) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_chunks) return;
  const int64_t s = t * NL_CHUNK, e = s + NL_CHUNK < n ? s + NL_CHUNK : n;
  int c = 0;
  for (int64_t k = s; k < e; ++k) c += d[k] == '\n';
  counts[t] = c;
}

__global__ void __launch_bounds__(256) k_nl_positions(const uint8_t* __restrict__ d, int64_t n,
                                                      const int32_t* __restrict__ first, int64_t n_chunks,
                                                      int64_t* nl) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_chunks) return;
  const int64_t s = t * NL_CHUNK, e = s + NL_CHUNK < n ? s + NL_CHUNK : n;
  int64_t o = first[t];
  for (int64_t k = s; k < e; ++k)
    if (d[k] == '\n') nl[o++] = k;
}

// line j of the text: [start, end) without the trailing "\r\n" characters
__device__ __forceinline__ void line_range(const uint8_t* d, const int64_t* nl, int64_t n_nl, int64_t n, int64_t j,
                                           int64_t& s, int64_t& e) {
  s = j == 0 ? 0 : nl[j - 1] + 1;
  e = j < n_nl ? nl[j] : n;
  while (e > s && (d[e - 1] == '\r' || d[e - 1] == '\n')) --e;
}

__device__ __forceinline__ int base_symbol_code_dev(int c) {   // codes.h:39
  return (((c & 0x4) >> 1)) | (((c & 0x4) >> 2) ^ ((c & 0x2) >> 1));
}

__device__ __forceinline__ int quality_level_dev(int q) {   // read_quality.c:10-16
  return q <= 2 ? 0 : q <= 7 ? 1 : q <= 10 ? 2 : 3;
}

// Record r = lines 4r .. 4r+3, typed as the reference types FASTQ lines by position and first
// character (get_input_line_type_fastq, load_data.c:85-107): a header '@', a sequence line (not
// empty, not '#'), a '+' line, and a quality line that is either a quality string or, when it
// starts with '#', a comment (the read then has no quality).  len[r] = read length (colors after
// the primer in color mode), or -1 when the record is anything else (the host loader then decides,
// and reports the file as corrupt where the reference does), -2 for an empty read.
__global__ void __launch_bounds__(256) k_fastq_len(const IngestArgs a) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.n_reads) return;
  int64_t hs, he, ss, se, ps, pe, qs, qe;
  line_range(a.text, a.nl, a.n_nl, a.n_bytes, 4 * r, hs, he);
  line_range(a.text, a.nl, a.n_nl, a.n_bytes, 4 * r + 1, ss, se);
  line_range(a.text, a.nl, a.n_nl, a.n_bytes, 4 * r + 2, ps, pe);
  line_range(a.text, a.nl, a.n_nl, a.n_bytes, 4 * r + 3, qs, qe);
  const bool regular = he > hs && a.text[hs] == '@' && se > ss && a.text[ss] != '#' && pe > ps &&
                       a.text[ps] == '+' && qe > qs;
  int L = -1;
  if (regular) {
    L = (int)(se - ss) - (a.color ? 2 : 0);
    if (L <= 0) L = -2;
  }
  a.len[r] = L;
  a.header[r] = hs;
  a.hasq[r] = (uint8_t)(qe > qs && a.text[qs] != '#');
}

__global__ void __launch_bounds__(256) k_iota_i32(int32_t* v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = (int32_t)i;
}

// one thread per read: codes, quality levels, tag, color primer fields; err[r] = 1 when a quality
// character lies below the -b base
__global__ void __launch_bounds__(256) k_fastq_pack(const IngestArgs a) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.n_reads) return;
  const int L = a.len[r];
  int64_t ss, se, qs, qe;
  line_range(a.text, a.nl, a.n_nl, a.n_bytes, 4 * r + 1, ss, se);
  line_range(a.text, a.nl, a.n_nl, a.n_bytes, 4 * r + 3, qs, qe);
  const uint8_t* s = a.text + ss;
  uint8_t* cs = a.codes + r * a.code_stride;
  uint32_t w = 0;
  uint64_t tag = 0, ct = 0;
  const int off = a.color ? 2 : 0;
  for (int k = 0; k < L; ++k) {
    const int c = s[k + off];
    int code;
    if (a.color) code = c & 3;
    else code = c == 'A' ? 0 : c == 'C' ? 1 : c == 'G' ? 2 : c == 'T' ? 3 : 0;   // codes.c:28-44
    w |= (uint32_t)code << ((k & 15) * 2);
    if ((k & 15) == 15) { *(uint32_t*)(cs + (k >> 4) * 4) = w; w = 0; }
    ct = (ct << 2) | (uint64_t)(((code + 1) & 3) ^ 1);
    if (ct > tag) tag = ct;
  }
  if (L & 15) *(uint32_t*)(cs + (L >> 4) * 4) = w;
  a.tag[r] = tag;
  if (a.color) a.first_base[r] = (uint8_t)(base_symbol_code_dev(s[0]) ^ (s[1] & 3));
  int bad = 0;
  if (a.qual && !(qe > qs && a.text[qs] != '#')) {   // quality line read as a comment: levels 3
    uint8_t* qo = a.qual + r * a.code_stride;
    for (int64_t b = 0; b < a.code_stride; b += 4) *(uint32_t*)(qo + b) = 0xFFFFFFFFu;
    if (a.color) a.first_qual[r] = 3;
  } else if (a.qual) {
    const uint8_t* q = a.text + qs;
    const int qlen = (int)(qe - qs);
    const int lim = L < qlen ? L : qlen;
    uint8_t* qo = a.qual + r * a.code_stride;
    // q[k] = min(char - base, 93) for k in [color ? 1 : 0, lim), 0 elsewhere (reads_data.c:516-531);
    // color: levels of q[1..L] (q[L] is past the reference's buffer and taken as 0)
    auto qv = [&](int k) -> int {
      if (k < (a.color ? 1 : 0) || k >= lim) return 0;
      const int v = (int)q[k] - a.qual_base;
      if (v < 0) bad = 1;
      return v < 93 ? v : 93;
    };
    uint32_t qw = 0;
    for (int k = 0; k < L; ++k) {
      const int lv = quality_level_dev(qv(k + (a.color ? 1 : 0)));
      qw |= (uint32_t)lv << ((k & 15) * 2);
      if ((k & 15) == 15) { *(uint32_t*)(qo + (k >> 4) * 4) = qw; qw = 0; }
    }
    if (L & 15) *(uint32_t*)(qo + (L >> 4) * 4) = qw;
    if (a.color) {
      const int v0 = (int)q[0] - a.qual_base;
      if (v0 < 0) bad = 1;
      a.first_qual[r] = (uint8_t)quality_level_dev(v0 < 93 ? v0 : 93);
    }
  }
  a.err[r] = (uint8_t)bad;
}

static dim3 blocks_of(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

hipError_t launch_nl_count(const uint8_t* d, int64_t n, int32_t* counts, int64_t n_chunks, hipStream_t st) {
  if (n_chunks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_nl_count, blocks_of(n_chunks), dim3(256), 0, st, d, n, counts, n_chunks);
  return hipGetLastError();
}

hipError_t launch_nl_positions(const uint8_t* d, int64_t n, const int32_t* first, int64_t n_chunks, int64_t* nl,
                               hipStream_t st) {
  if (n_chunks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_nl_positions, blocks_of(n_chunks), dim3(256), 0, st, d, n, first, n_chunks, nl);
  return hipGetLastError();
}

hipError_t launch_fastq_len(const IngestArgs& a, hipStream_t st) {
  if (a.n_reads <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fastq_len, blocks_of(a.n_reads), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_fastq_pack(const IngestArgs& a, hipStream_t st) {
  if (a.n_reads <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fastq_pack, blocks_of(a.n_reads), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_iota_i32(int32_t* v, int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_iota_i32, blocks_of(n), dim3(256), 0, st, v, n);
  return hipGetLastError();
}

int64_t nl_chunks(int64_t n) { return (n + NL_CHUNK - 1) / NL_CHUNK; }

// stable LSD radix sort of (tag, read index) pairs; the order ends in idx.Current()
hipError_t sort_tags(uint64_t* k0, uint64_t* k1, int32_t* v0, int32_t* v1, int64_t n, void* tmp, size_t& tmp_bytes,
                     int& result_sel, hipStream_t st) {
  rocprim::double_buffer<uint64_t> kb(k0, k1);
  rocprim::double_buffer<int32_t> vb(v0, v1);
  hipError_t e = rocprim::radix_sort_pairs(tmp, tmp_bytes, kb, vb, (size_t)n, 0, 64, st);
  result_sel = vb.current() == v0 ? 0 : 1;
  return e;
}

}  // namespace storm
