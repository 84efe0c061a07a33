// k_replay.hip -- per-read hit map (stage d tail) for gfx950.
//
// Semantics: the sequential part of process_read (run.c:660-697) and hit_map__update
// (hit_map.c:50-144).  For one read, one reference sequence and one strand the running
// threshold T starts at -t; the records (filter passes, ascending diagonal) are visited in order,
// a record is accepted when S >= T and score > T (the early exit of alignment.c:613-616), it
// updates the four slots of the read with the duplicate / ordered-insert rules, and T rises to
// slot[3].score when that is larger.  One thread owns one read; the slots persist across
// reference sequences and strands in HBM.
#include "storm_kernels.h"

namespace storm {

__global__ __launch_bounds__(128) void k_replay(ReplayArgs a) {
  const int64_t lr = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (lr >= a.n_reads) return;
  const int64_t read = a.read_begin + lr;
  Slot s[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) s[k] = a.slots[read * 4 + k];
  const int L = a.len[read];
  bool dirty = false;
  for (int strand = 0; strand < 2; ++strand) {
    if (!((a.sense_mask >> strand) & 1)) continue;
    const int64_t b = 2 * lr + strand;
    int T = a.t;
    for (int64_t r = a.rec_off[b]; r < a.rec_off[b + 1]; ++r) {
      const AlignRec x = a.rec[r];
      if (!(x.S >= T && x.score > T)) continue;
      const int score = x.score;
      const int ref_start = x.win + x.best_j0, ref_end = x.win + x.best_j;
      const int read_start = x.best_i0, read_end = L - 1;
      int k;
      bool drop = false;
      for (k = 0; k < 4 && s[k].ref_id != -1; ++k) {
        if (s[k].ref_id == a.ref_id &&
            ((s[k].ref_start == ref_start && s[k].read_start == read_start) ||
             (s[k].ref_end == ref_end && s[k].read_end == read_end))) {
          if (score > s[k].score) break;
          drop = true;
          break;
        }
      }
      if (!drop) {
        if (k == 4) k = 3;
        int i = k;
        while (i >= 0 && score > s[i].score) --i;
        if (i != 3) {
          const int32_t cell = s[k].cell;
          for (int j = k; j > i + 1; --j) s[j] = s[j - 1];
          Slot n;
          n.cell = cell;
          n.score = score; n.ref_id = a.ref_id; n.ref_start = ref_start; n.ref_end = ref_end;
          n.read_start = read_start; n.read_end = read_end; n.sense = strand ? -1 : 1;
          n.pass = a.pass; n.rec = (int32_t)r; n.tb_len = x.tb_len;
          s[i + 1] = n;
          dirty = true;
        }
      }
      if (s[3].score > T) T = s[3].score;
    }
  }
  if (dirty) {
#pragma unroll
    for (int k = 0; k < 4; ++k) a.slots[read * 4 + k] = s[k];
  }
}

__global__ __launch_bounds__(256) void k_slots_init(Slot* slots, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  Slot x;
  x.score = MIN_SCORE; x.ref_id = -1; x.ref_start = -1; x.ref_end = -1;
  x.read_start = -1; x.read_end = -1; x.sense = 0; x.pass = -1; x.rec = -1; x.tb_len = 0;
  x.cell = (int32_t)(t & 3);
  slots[t] = x;
}

// Tracebacks the pass inserted (ranks < max_rank: a slot never moves back up, so later ranks are
// never written out) go from the pass's record buffer to the read's cell store; the slot then
// names its cell (the four slots of a read hold a permutation of the four cells, so a cell is
// overwritten only when its previous entry has been evicted).
__global__ __launch_bounds__(256) void k_store_cells(Slot* slots, int64_t slot_begin, int64_t n_slots, int pass,
                                                     int max_rank, const uint8_t* tb, uint8_t* cells, const int32_t* ids) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_slots) return;
  const int64_t g = slot_begin + t;
  if ((int)(g & 3) >= max_rank) return;
  Slot s = slots[g];
  if (s.ref_id == -1 || s.pass != pass) return;
  const int64_t id = ids ? (int64_t)ids[g >> 2] : (g >> 2);
  const uint4* src = (const uint4*)(tb + (int64_t)s.rec * TBB);
  uint4* dst = (uint4*)(cells + (id * 4 + s.cell) * (int64_t)TBB);
#pragma unroll
  for (int w = 0; w < TBB / 16; ++w) dst[w] = src[w];
  s.pass = PASS_CELL;
  s.rec = s.cell;
  slots[g] = s;
}

hipError_t launch_store_cells(Slot* slots, int64_t slot_begin, int64_t n_slots, int pass, int max_rank, const uint8_t* tb,
                              uint8_t* cells, const int32_t* ids, hipStream_t st) {
  if (n_slots == 0) return hipSuccess;
  hipLaunchKernelGGL(k_store_cells, dim3((unsigned)((n_slots + 255) / 256)), dim3(256), 0, st, slots, slot_begin, n_slots,
                     pass, max_rank, tb, cells, ids);
  return hipGetLastError();
}

static_assert(TBB % 16 == 0, "cells are copied in 16-byte units");

// reads ids[0..k) of a batch (strand planes, quality planes, length, the four slots) into a
// contiguous subset, for the staged pass over the reads the fused kernel handed back
__global__ __launch_bounds__(256) void k_subset_gather(const uint32_t* planes, const uint32_t* qplanes, const int32_t* len,
                                                       const Slot* slots, const int32_t* ids, int64_t k, uint32_t* sp,
                                                       uint32_t* sq, int32_t* sl, Slot* ss) {
  constexpr int PW = 2 * 2 * RW, QWS = 2 * QW;   // words per read: both strands
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= k * (PW + QWS)) return;
  const int64_t b = t / (PW + QWS);
  const int w = (int)(t % (PW + QWS));
  const int64_t src = ids[b];
  if (w < PW) sp[b * PW + w] = planes[src * PW + w];
  else sq[b * QWS + (w - PW)] = qplanes[src * QWS + (w - PW)];
  if (w < 4) ss[b * 4 + w] = slots[src * 4 + w];
  if (w == 0) sl[b] = len[src];
}

__global__ __launch_bounds__(256) void k_subset_scatter(const Slot* ss, const int32_t* ids, int64_t k, Slot* slots) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 4 * k) return;
  slots[(int64_t)ids[t >> 2] * 4 + (t & 3)] = ss[t];
}

hipError_t launch_subset_gather(const uint32_t* planes, const uint32_t* qplanes, const int32_t* len, const Slot* slots,
                                const int32_t* ids, int64_t k, uint32_t* sp, uint32_t* sq, int32_t* sl, Slot* ss,
                                hipStream_t st) {
  if (k == 0) return hipSuccess;
  const int64_t total = k * (2 * 2 * RW + 2 * QW);
  hipLaunchKernelGGL(k_subset_gather, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, planes, qplanes, len, slots,
                     ids, k, sp, sq, sl, ss);
  return hipGetLastError();
}

hipError_t launch_subset_scatter(const Slot* ss, const int32_t* ids, int64_t k, Slot* slots, hipStream_t st) {
  if (k == 0) return hipSuccess;
  hipLaunchKernelGGL(k_subset_scatter, dim3((unsigned)((4 * k + 255) / 256)), dim3(256), 0, st, ss, ids, k, slots);
  return hipGetLastError();
}

// rows of a byte matrix (stride bytes per row) into the order given: dst row i = src row order[i]
__global__ __launch_bounds__(256) void k_permute_bytes(const uint8_t* src, int64_t stride, const int32_t* order, int64_t n,
                                                       uint8_t* dst) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * stride) return;
  const int64_t i = t / stride, b = t % stride;
  dst[t] = src[(int64_t)order[i] * stride + b];
}

__global__ __launch_bounds__(256) void k_permute_i32(const int32_t* src, const int32_t* order, int64_t n, int32_t* dst) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[t] = src[order[t]];
}

// the per-read slots (device order) as four int32 arrays in the caller's order: entry order[i]*4+k
__global__ __launch_bounds__(256) void k_hits_out(const Slot* slots, const int32_t* order, int64_t n, int32_t* score,
                                                  int32_t* ref_id, int32_t* ref_start, int32_t* sense) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 4 * n) return;
  const int64_t o = (order ? (int64_t)order[t >> 2] : (t >> 2)) * 4 + (t & 3);
  const Slot s = slots[t];
  score[o] = s.score;
  ref_id[o] = s.ref_id;
  ref_start[o] = s.ref_start;
  sense[o] = s.sense;
}

hipError_t launch_permute_bytes(const uint8_t* src, int64_t stride, const int32_t* order, int64_t n, uint8_t* dst,
                                hipStream_t st) {
  if (n == 0 || stride == 0) return hipSuccess;
  hipLaunchKernelGGL(k_permute_bytes, dim3((unsigned)((n * stride + 255) / 256)), dim3(256), 0, st, src, stride, order, n,
                     dst);
  return hipGetLastError();
}

hipError_t launch_permute_i32(const int32_t* src, const int32_t* order, int64_t n, int32_t* dst, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_permute_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src, order, n, dst);
  return hipGetLastError();
}

hipError_t launch_hits_out(const Slot* slots, const int32_t* order, int64_t n, int32_t* score, int32_t* ref_id,
                           int32_t* ref_start, int32_t* sense, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_hits_out, dim3((unsigned)((4 * n + 255) / 256)), dim3(256), 0, st, slots, order, n, score, ref_id,
                     ref_start, sense);
  return hipGetLastError();
}

hipError_t launch_replay(const ReplayArgs& a, hipStream_t st) {
  if (a.n_reads == 0) return hipSuccess;
  hipLaunchKernelGGL(k_replay, dim3((unsigned)((a.n_reads + 127) / 128)), dim3(128), 0, st, a);
  return hipGetLastError();
}
hipError_t launch_slots_init(Slot* slots, int64_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_slots_init, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, slots, n);
  return hipGetLastError();
}


// The whole hit map in the caller's order: one thread per slot writes its record and, for a hit
// at rank < max_rank, copies its traceback cell (16 bytes at a time) to the dense output.
__global__ __launch_bounds__(256) void k_hit_records(const Slot* slots, const int32_t* order, int64_t r0, int64_t r1,
                                                     int max_rank, const uint8_t* cells, HitRec* out, uint8_t* tb_out,
                                                     int tb_stride, int* mapped) {
  const int64_t t = 4 * r0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 4 * r1) return;
  const int k = (int)(t & 3);
  const int64_t row = order ? (int64_t)order[t >> 2] : (t >> 2);
  const Slot s = slots[t];
  if (mapped) {   // reads whose best slot holds a hit: one add per wavefront, by its first live lane
    const unsigned long long hit = __ballot(k == 0 && s.ref_id != -1);
    const unsigned long long live = __ballot(1);
    if ((int)(threadIdx.x & 63) == __ffsll((long long)live) - 1 && hit) atomicAdd(mapped, (int)__popcll(hit));
  }
  HitRec h;
  h.score = s.score; h.ref_id = s.ref_id; h.ref_start = s.ref_start; h.ref_end = s.ref_end;
  h.read_start = s.read_start; h.read_end = s.read_end; h.sense = s.sense;
  h.tb_len = s.ref_id == -1 ? 0 : s.tb_len;
  out[row * 4 + k] = h;
  if (tb_out && k < max_rank) {
    // the traceback's (tb_len + 1) / 2 bytes, the rest of the record zero (all of it for an empty
    // slot), so the output depends on the hit map alone
    const bool hit = s.ref_id != -1 && s.pass == PASS_CELL;
    const int nb = hit ? (h.tb_len + 1) / 2 : 0;
    const uint4* src = (const uint4*)(cells + ((t & ~3LL) + s.cell) * (int64_t)TBB);
    uint4* dst = (uint4*)(tb_out + (row * max_rank + k) * (int64_t)tb_stride);
#pragma unroll
    for (int w = 0; w < TBB / 16; ++w) {
      if (16 * w >= tb_stride) break;
      uint4 x = make_uint4(0, 0, 0, 0);
      if (16 * w < nb) {
        x = src[w];
        const int keep = nb - 16 * w;   // bytes of this word that belong to the traceback
        if (keep < 16) {
          uint32_t* xs = (uint32_t*)&x;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int kb = keep - 4 * q;
            xs[q] = kb >= 4 ? xs[q] : kb <= 0 ? 0u : (xs[q] & ((1u << (8 * kb)) - 1u));
          }
        }
      }
      dst[w] = x;
    }
  }
}

hipError_t launch_hit_records(const Slot* slots, const int32_t* order, int64_t r0, int64_t r1, int max_rank,
                              const uint8_t* cells, HitRec* out, uint8_t* tb_out, int tb_stride, int* mapped,
                              hipStream_t st) {
  if (r1 <= r0) return hipSuccess;
  if (tb_stride <= 0 || tb_stride > TBB || tb_stride % 16) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_hit_records, dim3((unsigned)((4 * (r1 - r0) + 255) / 256)), dim3(256), 0, st, slots, order, r0,
                     r1, max_rank, cells, out, tb_out, tb_stride, mapped);
  return hipGetLastError();
}

}  // namespace storm
