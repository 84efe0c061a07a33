// k_serve.hip -- batch assembly of the serving loop: the reads of a batch are gathered from the
// device regions of the requests they belong to (codes, quality levels, lengths) into one packed
// batch, and the batch's hit-map slots are scattered back to each request's result array.
#include "storm_kernels.h"

namespace storm {

// one thread per (read, 32-bit word): words [0, SERVE_CS/4) of codes, then the same of levels
__global__ __launch_bounds__(256) void k_serve_gather(ServeGatherArgs a) {
  constexpr int WPR = SERVE_CS / 4;   // words per plane per read
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n * 2 * WPR) return;
  const int64_t b = t / (2 * WPR);
  const int w = (int)(t % (2 * WPR));
  const ServeRow rw = a.rows[b];
  const ServeRegion& g = a.regions[rw.region];
  const int64_t row = rw.row;
  if (w < WPR) {
    const uint8_t* src = g.codes + row * g.code_stride;
    uint32_t v = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int off = 4 * w + k;
      if (off < g.code_stride) v |= (uint32_t)src[off] << (8 * k);
    }
    ((uint32_t*)(a.codes + b * SERVE_CS))[w] = v;
    if (w == 0) {
      a.len[b] = g.len[row];
      // a read without levels scores as level 3 everywhere (alignment.c:151-162), including the
      // primer-side color that the reverse strand puts last (read_quality.c:47-64)
      a.first_qual[b] = g.qual ? (g.first_qual ? g.first_qual[row] : (uint8_t)0) : (uint8_t)3;
    }
  } else {
    const int wq = w - WPR;
    uint32_t v = 0xffffffffu;
    if (g.qual) {
      const uint8_t* src = g.qual + row * g.qual_stride;
      v = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int off = 4 * wq + k;
        if (off < g.qual_stride) v |= (uint32_t)src[off] << (8 * k);
      }
    }
    ((uint32_t*)(a.qual + b * SERVE_CS))[wq] = v;
  }
}

// one thread per (read, slot)
__global__ __launch_bounds__(256) void k_serve_scatter(const Slot* slots, const ServeRow* rows, int64_t n,
                                                       const ServeRegion* regions) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 4 * n) return;
  const int64_t b = t >> 2;
  const int k = (int)(t & 3);
  const ServeRow rw = rows[b];
  const Slot s = slots[t];
  int4 v;
  v.x = s.score;
  v.y = s.ref_id;
  v.z = s.ref_start;
  v.w = s.sense;
  ((int4*)regions[rw.region].hits)[(int64_t)rw.row * 4 + k] = v;
}

hipError_t launch_serve_gather(const ServeGatherArgs& a, hipStream_t st) {
  if (a.n == 0) return hipSuccess;
  const int64_t threads = a.n * 2 * (SERVE_CS / 4);
  hipLaunchKernelGGL(k_serve_gather, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_serve_scatter(const Slot* slots, const ServeRow* rows, int64_t n, const ServeRegion* regions,
                                hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_serve_scatter, dim3((unsigned)((4 * n + 255) / 256)), dim3(256), 0, st, slots, rows, n,
                     regions);
  return hipGetLastError();
}

}  // namespace storm
