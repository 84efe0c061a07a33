// k_index.hip -- genome bit planes and the CSR seed index built on the device.
//
// index__build_reference (index.c:19-195): every window start q in [0, size - span] outside the
// skipped ranges of the masked regions gets the key of the seed applied at q
// (seed__apply_to_compressed, seed.c:203-214); first[k]..first[k+1] delimit the positions of key k
// in ascending order (index.c:157-188), keys erased by the repeat filter own no positions.
//
// Device formulation: the keys of all window starts are sorted with their positions by a stable
// LSD radix sort (skipped windows carry the key 4^w and sort last), so each key's positions come
// out ascending; the run of each key in the sorted order gives its count (start and end of the
// run, no atomics), an exclusive scan of the counts left by the repeat filter gives `first`, and
// every sorted position is written at first[key] + its rank inside the run.
#include <hip/hip_runtime.h>
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "storm_kernels.h"

namespace storm {

// packed 2-bit codes (4 per byte, code n in bits 2(n & 3) of byte n >> 2, codes.h:125-130) to
// one code per byte
__global__ void __launch_bounds__(256) k_unpack_codes(const uint8_t* __restrict__ packed, int64_t size,
                                                      uint8_t* code) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= size) return;
  code[i] = (packed[i >> 2] >> ((i & 3) << 1)) & 3;
}

// One thread per 32-bit plane word: code bit 0, code bit 1, mask bit of the codes
// [32 w - GPAD, 32 w - GPAD + 32).  Codes outside [0, size) are 0 and masked.
__global__ void __launch_bounds__(256) k_genome_planes(const uint8_t* __restrict__ code,
                                                       const uint8_t* __restrict__ mask, int64_t size,
                                                       int64_t nw, uint2* lh, uint32_t* mk) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nw) return;
  const int64_t p0 = w * 32 - GPAD;
  uint32_t l = 0, h = 0, m = 0;
#pragma unroll 8
  for (int b = 0; b < 32; ++b) {
    const int64_t p = p0 + b;
    if (p >= 0 && p < size) {
      const uint32_t c = code[p];
      l |= (c & 1u) << b;
      h |= ((c >> 1) & 1u) << b;
      if (mask && mask[p]) m |= 1u << b;
    } else {
      m |= 1u << b;
    }
  }
  lh[w] = make_uint2(l, h);
  mk[w] = m;
}

// One thread per window start q < map_size: the seed key, or ks (= 4^w) when q is skipped.
// regions: masked regions [start, end) sorted and disjoint.  The reference's scan
// (index.c:66-72) jumps from start - span to end - 1 and indexes end - 1, so q is skipped iff
// start - span <= q <= end - 2 for some region; the first region with end >= q + 2 is the only
// candidate (later ones start after its end).
__global__ void __launch_bounds__(256) k_index_keys(const uint8_t* __restrict__ code, int64_t map_size,
                                                    const int2* __restrict__ regions, int n_regions, SeedDev seed,
                                                    uint32_t ks, uint32_t* keys, int32_t* vals) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= map_size) return;
  int lo = 0, hi = n_regions;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if ((int64_t)regions[mid].y < q + 2) lo = mid + 1;
    else hi = mid;
  }
  bool skip = lo < n_regions && (int64_t)regions[lo].x - seed.span <= q;
  uint32_t key = 0;
  for (int k = 0; k < seed.weight; ++k) key = (key << 2) | code[q + seed.offs[k]];
  keys[q] = skip ? ks : key;
  vals[q] = (int32_t)q;
}

// Runs of equal keys in the sorted order: st[k] = first index, en[k] = one past the last.
__global__ void __launch_bounds__(256) k_index_runs(const uint32_t* __restrict__ keys, int64_t n, int32_t* st,
                                                    int32_t* en) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const uint32_t k = keys[e];
  if (e == 0 || keys[e - 1] != k) st[k] = (int32_t)e;
  if (e == n - 1 || keys[e + 1] != k) en[k] = (int32_t)(e + 1);
}

// counts[k] = en[k] - st[k] for k < ks (in place over en), 0 for the sentinel slot ks
__global__ void __launch_bounds__(256) k_index_counts(const int32_t* __restrict__ st, int32_t* en, int64_t ks) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k > ks) return;
  en[k] = k < ks ? en[k] - st[k] : 0;
}

// counts of erased keys to 0 (erased: one byte per key)
__global__ void __launch_bounds__(256) k_index_erase(const uint8_t* __restrict__ erased, int32_t* counts,
                                                     int64_t ks) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ks) return;
  if (erased[k]) counts[k] = 0;
}

// map[first[k] + e - st[k]] = position of sorted entry e, for kept keys
__global__ void __launch_bounds__(256) k_index_scatter(const uint32_t* __restrict__ keys,
                                                       const int32_t* __restrict__ vals, int64_t n,
                                                       const int32_t* __restrict__ st,
                                                       const int32_t* __restrict__ first,
                                                       const uint8_t* __restrict__ erased, uint32_t ks,
                                                       int32_t* map) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const uint32_t k = keys[e];
  if (k >= ks) return;
  if (erased && erased[k]) return;
  map[first[k] + (int32_t)(e - st[k])] = vals[e];
}

static dim3 blocks_for(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

hipError_t launch_unpack_codes(const uint8_t* packed, int64_t size, uint8_t* code, hipStream_t st) {
  if (size <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_unpack_codes, blocks_for(size), dim3(256), 0, st, packed, size, code);
  return hipGetLastError();
}

hipError_t launch_genome_planes(const uint8_t* code, const uint8_t* mask, int64_t size, int64_t nw, uint2* lh,
                                uint32_t* mk, hipStream_t st) {
  if (nw <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_genome_planes, blocks_for(nw), dim3(256), 0, st, code, mask, size, nw, lh, mk);
  return hipGetLastError();
}

hipError_t launch_index_keys(const uint8_t* code, int64_t map_size, const int2* regions, int n_regions,
                             const SeedDev& seed, uint32_t ks, uint32_t* keys, int32_t* vals, hipStream_t st) {
  if (map_size <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_index_keys, blocks_for(map_size), dim3(256), 0, st, code, map_size, regions, n_regions, seed,
                     ks, keys, vals);
  return hipGetLastError();
}

hipError_t launch_index_runs(const uint32_t* keys, int64_t n, int32_t* st_, int32_t* en, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_index_runs, blocks_for(n), dim3(256), 0, st, keys, n, st_, en);
  return hipGetLastError();
}

hipError_t launch_index_counts(const int32_t* st_, int32_t* en, int64_t ks, hipStream_t st) {
  hipLaunchKernelGGL(k_index_counts, blocks_for(ks + 1), dim3(256), 0, st, st_, en, ks);
  return hipGetLastError();
}

hipError_t launch_index_erase(const uint8_t* erased, int32_t* counts, int64_t ks, hipStream_t st) {
  hipLaunchKernelGGL(k_index_erase, blocks_for(ks), dim3(256), 0, st, erased, counts, ks);
  return hipGetLastError();
}

hipError_t launch_index_scatter(const uint32_t* keys, const int32_t* vals, int64_t n, const int32_t* st_,
                                const int32_t* first, const uint8_t* erased, uint32_t ks, int32_t* map,
                                hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_index_scatter, blocks_for(n), dim3(256), 0, st, keys, vals, n, st_, first, erased, ks, map);
  return hipGetLastError();
}

// stable LSD radix sort of (key, position) pairs over bits [0, end_bit); result_sel = 1 when the
// sorted pairs end in (k1, v1), 0 when in (k0, v0)
hipError_t sort_key_pos_raw(uint32_t* k0, uint32_t* k1, int32_t* v0, int32_t* v1, int64_t n, int end_bit, void* tmp,
                            size_t& tmp_bytes, int& result_sel, hipStream_t st) {
  rocprim::double_buffer<uint32_t> kb(k0, k1);
  rocprim::double_buffer<int32_t> vb(v0, v1);
  hipError_t e = rocprim::radix_sort_pairs(tmp, tmp_bytes, kb, vb, (size_t)n, 0, (unsigned)end_bit, st);
  result_sel = kb.current() == k0 ? 0 : 1;
  return e;
}

hipError_t scan_exclusive_i32(const int32_t* in, int32_t* out, int64_t n, void* tmp, size_t& tmp_bytes,
                              hipStream_t st) {
  return rocprim::exclusive_scan(tmp, tmp_bytes, in, out, (int32_t)0, (size_t)n, rocprim::plus<int32_t>(), st);
}

}  // namespace storm
