// k_seed.hip -- read preparation, spaced-seed keys, index lookup and candidate diagonals
// (stages a + b) for gfx950.
//
// Semantics (run.c:555-747, seed.c:203-214, index.c:212-241): for a read strand, every seed s and
// read position rp give key = codes at the '#' offsets (first one most significant); the CSR
// bucket first[key]..first[key+1] of map[] holds the genome positions q; the candidates are the
// diagonals q - rp in ascending order without repeats, where the smallest one is dropped when it
// is <= -I (old_hit_pos starts at -I and is overwritten on every pop, run.c:617,648,697).
//
// Layout: one wavefront per read strand.  The keys and bucket sizes come from one pass, a
// prefix sum across the wave places every hit, the diagonals are bitonic-sorted in LDS and the
// keep rule is applied in the same pass.  A read strand with more hits than LDS_HITS is flagged
// and finished by the overflow path (device radix sort of its segment, then k_unique_segments).
#include "storm_kernels.h"

#include <cstring>

#include <rocprim/rocprim.hpp>

namespace storm {

constexpr int LDS_HITS = 2048;

// --- read preparation: forward and reverse-complement planes + quality levels -------------------
__global__ __launch_bounds__(256) void k_prep_reads(PrepArgs a) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 * a.n) return;
  const int64_t r = t >> 1;
  const int strand = (int)(t & 1);
  const int L = a.len[r];
  const uint8_t* src = a.codes + r * a.code_stride;
  uint32_t lo[RW], hi[RW], q[QW];
#pragma unroll
  for (int w = 0; w < RW; ++w) { lo[w] = 0; hi[w] = 0; }
#pragma unroll
  for (int w = 0; w < QW; ++w) q[w] = 0;
  for (int i = 0; i < L; ++i) {
    // forward: code i; reverse (codes.c:429-439): dest[L-1-k] = comp(src[k]) (color: no complement)
    const int k = strand ? (L - 1 - i) : i;
    int c = (src[k >> 2] >> ((k & 3) << 1)) & 3;
    if (strand && !a.color) c = (~c) & 3;
    lo[i >> 5] |= (uint32_t)(c & 1) << (i & 31);
    hi[i >> 5] |= (uint32_t)((c >> 1) & 1) << (i & 31);
    if (a.qual) {
      int lv;
      if (!strand) {
        const uint8_t* qs = a.qual + r * a.qual_stride;
        lv = (qs[i >> 2] >> ((i & 3) << 1)) & 3;
      } else {
        // read_quality.c:47-64: nucleotide reversed; color dest[L-1] = first, dest[i] = src[L-2-i]
        const uint8_t* qs = a.qual + r * a.qual_stride;
        if (!a.color) { const int s = L - 1 - i; lv = (qs[s >> 2] >> ((s & 3) << 1)) & 3; }
        else if (i == L - 1) lv = a.first_qual ? a.first_qual[r] : 0;
        else { const int s = L - 2 - i; lv = (qs[s >> 2] >> ((s & 3) << 1)) & 3; }
      }
      q[i >> 4] |= (uint32_t)lv << ((i & 15) << 1);
    }
  }
  uint32_t* dst = a.planes + t * (2 * RW);
#pragma unroll
  for (int w = 0; w < RW; ++w) { dst[w] = lo[w]; dst[RW + w] = hi[w]; }
  uint32_t* qd = a.qplanes + t * QW;
#pragma unroll
  for (int w = 0; w < QW; ++w) qd[w] = q[w];
}

// --- seeds -----------------------------------------------------------------------------------
// wave-wide inclusive scan (64 lanes)
__device__ __forceinline__ int64_t wave_scan_incl(int64_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int64_t n = __shfl_up(v, o, 64);
    if (lane >= o) v += n;
  }
  return v;
}

__global__ __launch_bounds__(64) void k_seed_count(SeedArgs a) {
  const int64_t rs = a.rs_begin + blockIdx.x;
  if (blockIdx.x >= a.n_rs) return;
  const int lane = threadIdx.x;
  int64_t cnt = 0;
  if ((a.sense_mask >> (rs & 1)) & 1) {
    const int L = a.r.len[rs >> 1];
    const uint32_t* pl = read_planes(a.r, rs);
    for (int s = 0; s < a.seeds.n; ++s) {
      const SeedDev& sd = a.seeds.s[s];
      const int np = seed_npos(sd, L);
      for (int k = lane; k < np; k += 64) {
        const int rp = sd.n_pos > 0 ? sd.pos[k] : k;
        const int key = seed_key(sd, pl, rp);
        cnt += sd.first[key + 1] - sd.first[key];
      }
    }
  }
  cnt = wave_scan_incl(cnt);
  if (lane == 63) a.seg_count[blockIdx.x] = cnt;
}

__global__ __launch_bounds__(64) void k_seed_gather(SeedArgs a) {
  __shared__ int32_t buf[LDS_HITS];
  const int64_t b = blockIdx.x;
  if (b >= a.n_rs) return;
  const int64_t rs = a.rs_begin + b;
  const int lane = threadIdx.x;
  const int64_t base = a.seg_off[b];
  const int64_t total = a.seg_off[b + 1] - base;
  if (total == 0) { if (lane == 0) { a.overflow[b] = 0; a.seg_count[b] = 0; } return; }
  const bool in_lds = total <= LDS_HITS;
  const int L = a.r.len[rs >> 1];
  const uint32_t* pl = read_planes(a.r, rs);
  int64_t run = 0;   // hits placed by earlier chunks
  for (int s = 0; s < a.seeds.n; ++s) {
    const SeedDev& sd = a.seeds.s[s];
    const int np = seed_npos(sd, L);
    for (int k0 = 0; k0 < np; k0 += 64) {
      const int k = k0 + lane;
      int rp = 0, f0 = 0, c = 0;
      if (k < np) {
        rp = sd.n_pos > 0 ? sd.pos[k] : k;
        const int key = seed_key(sd, pl, rp);
        f0 = sd.first[key];
        c = sd.first[key + 1] - f0;
      }
      const int64_t incl = wave_scan_incl(c);
      const int64_t off = run + incl - c;
      for (int e = 0; e < c; ++e) {
        const int32_t dg = sd.map[f0 + e] - rp;
        if (in_lds) buf[off + e] = dg;
        else a.hits[base + off + e] = dg;
      }
      run += __shfl(incl, 63, 64);
    }
  }
  if (!in_lds) {
    if (lane == 0) { a.overflow[b] = 1; atomicAdd(a.n_overflow, 1); }
    return;
  }
  // bitonic sort of the segment, padded to a power of two with INT_MAX
  int np2 = 1;
  while (np2 < total) np2 <<= 1;
  for (int e = (int)total + lane; e < np2; e += 64) buf[e] = 0x7fffffff;
  __syncthreads();
  for (int kk = 2; kk <= np2; kk <<= 1) {
    for (int j = kk >> 1; j > 0; j >>= 1) {
      for (int e = lane; e < np2; e += 64) {
        const int x = e ^ j;
        if (x > e) {
          const int32_t u = buf[e], v = buf[x];
          const bool up = (e & kk) == 0;
          if ((u > v) == up) { buf[e] = v; buf[x] = u; }
        }
      }
      __syncthreads();
    }
  }
  // keep rule and in-order compaction
  int64_t placed = 0;
  for (int e0 = 0; e0 < total; e0 += 64) {
    const int e = e0 + lane;
    int keep = 0;
    int32_t v = 0;
    if (e < total) {
      v = buf[e];
      keep = (e == 0) ? (v > -a.I) : (v > buf[e - 1]);
    }
    const int64_t incl = wave_scan_incl(keep);
    if (keep) a.kept[base + placed + incl - 1] = v;
    placed += __shfl(incl, 63, 64);
  }
  if (lane == 0) { a.overflow[b] = 0; a.seg_count[b] = placed; }
}

// overflow segments, after a device radix sort in `hits`: keep rule into `kept`
__global__ __launch_bounds__(64) void k_unique_segments(SeedArgs a) {
  const int64_t b = blockIdx.x;
  if (b >= a.n_rs || !a.overflow[b]) return;
  const int lane = threadIdx.x;
  const int64_t base = a.seg_off[b];
  const int64_t total = a.seg_off[b + 1] - base;
  int64_t placed = 0;
  for (int64_t e0 = 0; e0 < total; e0 += 64) {
    const int64_t e = e0 + lane;
    int keep = 0;
    int32_t v = 0;
    if (e < total) {
      v = a.hits[base + e];
      keep = (e == 0) ? (v > -a.I) : (v > a.hits[base + e - 1]);
    }
    const int64_t incl = wave_scan_incl(keep);
    if (keep) a.kept[base + placed + incl - 1] = v;
    placed += __shfl(incl, 63, 64);
  }
  if (lane == 0) a.seg_count[b] = placed;
}

// dense candidate list: (read strand, window start) in read-strand then diagonal order
__global__ __launch_bounds__(64) void k_compact_candidates(CompactArgs a) {
  const int64_t b = blockIdx.x;
  if (b >= a.n_rs) return;
  const int64_t src = a.seg_off[b], dst = a.cand_off[b], n = a.seg_count[b];
  for (int64_t e = threadIdx.x; e < n; e += 64) {
    const int32_t dg = a.kept[src + e];
    a.cand_rs[dst + e] = (int32_t)(a.rs_begin + b);
    a.cand_diag[dst + e] = dg;
    a.cand_pos[dst + e] = dg - a.d;
  }
}

// records = candidates that passed the filter, in the same order; rec_off[b] = first record of
// read strand b (exclusive scan of the pass flags read at each segment start)
__global__ __launch_bounds__(256) void k_make_records(RecArgs a) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < a.n_cand && a.pass[t]) {
    const int64_t o = a.pass_scan[t];
    a.rec_rs[o] = a.cand_rs[t];
    a.rec_diag[o] = a.cand_diag[t];
  }
  if (t <= a.n_rs) a.rec_off[t] = a.pass_scan[a.cand_off[t]];
}


__global__ __launch_bounds__(256) void k_overflow_ranges(const int64_t* seg_off, const uint8_t* overflow, int64_t n_rs,
                                                        int64_t* begin, int64_t* end) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n_rs) return;
  begin[b] = seg_off[b];
  end[b] = overflow[b] ? seg_off[b + 1] : seg_off[b];
}

hipError_t launch_overflow_ranges(const int64_t* seg_off, const uint8_t* overflow, int64_t n_rs, int64_t* begin,
                                  int64_t* end, hipStream_t st) {
  if (n_rs == 0) return hipSuccess;
  hipLaunchKernelGGL(k_overflow_ranges, dim3((unsigned)((n_rs + 255) / 256)), dim3(256), 0, st, seg_off, overflow, n_rs,
                     begin, end);
  return hipGetLastError();
}

hipError_t launch_prep_reads(const PrepArgs& a, hipStream_t st) {
  if (a.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_prep_reads, dim3((unsigned)((2 * a.n + 255) / 256)), dim3(256), 0, st, a);
  return hipGetLastError();
}
hipError_t launch_seed_count(const SeedArgs& a, hipStream_t st) {
  if (a.n_rs == 0) return hipSuccess;
  hipLaunchKernelGGL(k_seed_count, dim3((unsigned)a.n_rs), dim3(64), 0, st, a);
  return hipGetLastError();
}
hipError_t launch_seed_gather(const SeedArgs& a, hipStream_t st) {
  if (a.n_rs == 0) return hipSuccess;
  hipLaunchKernelGGL(k_seed_gather, dim3((unsigned)a.n_rs), dim3(64), 0, st, a);
  return hipGetLastError();
}
hipError_t launch_unique_segments(const SeedArgs& a, hipStream_t st) {
  if (a.n_rs == 0) return hipSuccess;
  hipLaunchKernelGGL(k_unique_segments, dim3((unsigned)a.n_rs), dim3(64), 0, st, a);
  return hipGetLastError();
}
hipError_t launch_compact_candidates(const CompactArgs& a, hipStream_t st) {
  if (a.n_rs == 0) return hipSuccess;
  hipLaunchKernelGGL(k_compact_candidates, dim3((unsigned)a.n_rs), dim3(64), 0, st, a);
  return hipGetLastError();
}
hipError_t launch_make_records(const RecArgs& a, hipStream_t st) {
  const int64_t n = (a.n_cand > a.n_rs + 1) ? a.n_cand : a.n_rs + 1;
  hipLaunchKernelGGL(k_make_records, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t scan_exclusive_i64(const int64_t* in, int64_t* out, int64_t n, void* tmp, size_t& tmp_bytes, hipStream_t st) {
  return rocprim::exclusive_scan(tmp, tmp_bytes, in, out, (int64_t)0, (size_t)n, rocprim::plus<int64_t>(), st);
}

namespace {
struct WidenU8 {
  __host__ __device__ int64_t operator()(uint8_t x) const { return (int64_t)x; }
};
}  // namespace

hipError_t scan_exclusive_u8(const uint8_t* in, int64_t* out, int64_t n, void* tmp, size_t& tmp_bytes, hipStream_t st) {
  rocprim::transform_iterator<const uint8_t*, WidenU8, int64_t> it(in, WidenU8());
  return rocprim::exclusive_scan(tmp, tmp_bytes, it, out, (int64_t)0, (size_t)n, rocprim::plus<int64_t>(), st);
}

// the raw diagonals of each overflowing read strand, ascending (segments [begin, end) of `in`)
hipError_t segsort_i32(const int32_t* in, int32_t* out, int64_t n_items, int64_t n_seg, const int64_t* begin,
                       const int64_t* end, void* tmp, size_t& tmp_bytes, hipStream_t st) {
  return rocprim::segmented_radix_sort_keys(tmp, tmp_bytes, in, out, (unsigned)n_items, (unsigned)n_seg, begin, end, 0,
                                            32, st);
}

}  // namespace storm
