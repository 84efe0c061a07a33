// k_fused.hip -- persistent fused seed -> lookup -> filter -> DP -> hit map kernel (batched path).
//
// Same semantics as the staged kernels (k_seed, k_filter, k_align, k_replay; run.c:555-747,
// hit_map.c:50-144), in one launch.  Workgroups are single wavefronts that pull tasks from a
// device counter until the batch is exhausted (every wave reaches the exit).  A task is a group
// of G consecutive reads (both strands): its candidates, filter flags and DP records stay in
// LDS, so a group of reads keeps all 64 lanes busy even though one read strand has only a few
// candidates.  Reads of any length share a task; one launch serves one band class (W, I).
// A read whose raw hits, candidates or filter passes exceed the LDS capacities is flagged in
// `overflow` and left untouched; the host sends it through the staged path.
#include "dev_align.h"
#include "dev_filter.h"
#include "dev_seed.h"
#include "storm_kernels.h"

namespace storm {

constexpr int F_HCAP = 1024;   // raw hits of one read strand
constexpr int F_CCAP = 768;    // candidates of one task
constexpr int F_PCAP = FUSED_PCAP;   // filter passes (DP records) of one task
constexpr int F_GMAX = 64;     // reads per task


template <int W, int IM, bool COLOR>
__global__ __launch_bounds__(64) void k_fused(FusedArgs a) {
  __shared__ int32_t hbuf[F_HCAP];
  __shared__ int32_t cdiag[F_CCAP];
  __shared__ uint8_t crs[F_CCAP];
  __shared__ uint8_t cpass[F_CCAP];
  __shared__ int16_t crec[F_CCAP];
  __shared__ int16_t cstart[2 * F_GMAX + 1];
  __shared__ uint8_t rovf[F_GMAX];
  __shared__ int32_t rawh[2 * F_GMAX];
  __shared__ AlignRec recs[F_PCAP];
  __shared__ int16_t rec2cand[F_PCAP];
  __shared__ int32_t task_sh;

  const int lane = threadIdx.x;
  const int n = *a.n_reads;
  const int G = a.group;
  const int I = a.sc.I;
  const FilterConst fk = filter_const(a.sc.fmatch, a.sc.fmismatch, a.sc.fopen, a.sc.fext, a.sc.fthr);
  uint8_t* scratch = a.scratch + (int64_t)blockIdx.x * 64 * a.scratch_stride + (int64_t)lane * a.scratch_stride;
  uint8_t* tbbuf = a.tbbuf + (int64_t)blockIdx.x * F_PCAP * TBB;
  unsigned long long st_cand = 0, st_pass = 0, st_acc = 0, st_hits = 0;

  for (;;) {
    if (lane == 0) task_sh = atomicAdd(a.task_counter, 1);
    __syncthreads();
    const int task = task_sh;
    __syncthreads();
    const int64_t base = (int64_t)task * G;
    if (base >= n) break;
    const int nr = (n - base) < G ? (int)(n - base) : G;
    for (int j = lane; j < nr; j += 64) rovf[j] = 0;
    __syncthreads();

    // 1. candidates of every read strand of the task (stages a + b)
    int nc = 0;
    for (int j = 0; j < 2 * nr; ++j) {
      if (lane == 0) { cstart[j] = (int16_t)nc; rawh[j] = 0; }
      const bool wanted = ((a.sense_mask >> (j & 1)) & 1) && !rovf[j >> 1];
      if (!wanted) continue;
      const int64_t rs = 2 * base + j;
      int raw = 0;
      const int cnt = wave_candidates(a.r, a.seeds, rs, I, hbuf, F_HCAP, cdiag + nc, F_CCAP - nc, &raw);
      if (lane == 0) rawh[j] = raw;
      if (cnt < 0) {
        if (lane == 0) rovf[j >> 1] = 1;
        __syncthreads();
        continue;
      }
      for (int e = lane; e < cnt; e += 64) crs[nc + e] = (uint8_t)j;
      nc += cnt;
    }
    if (lane == 0) cstart[2 * nr] = (int16_t)nc;
    __syncthreads();

    // 2. filter (stage c): two candidates per lane
    for (int b0 = 0; b0 < nc; b0 += 128) {
      const int c0 = b0 + 2 * lane;
      if (c0 < nc) {
        const int c1 = (c0 + 1 < nc) ? c0 + 1 : c0;
        const int64_t rsa = 2 * base + crs[c0], rsb = 2 * base + crs[c1];
        const int La = a.r.len[rsa >> 1], Lb = a.r.len[rsb >> 1];
        int sa, sb;
        filter_pair<W>(a.g, read_planes(a.r, rsa), read_planes(a.r, rsb), La, Lb,
                       (int64_t)cdiag[c0] - a.sc.d + GPAD, (int64_t)cdiag[c1] - a.sc.d + GPAD, fk, sa, sb);
        cpass[c0] = (uint8_t)(sa > fk.thr);
        if (c1 != c0) cpass[c1] = (uint8_t)(sb > fk.thr);
      }
    }
    __syncthreads();

    // 3. DP + traceback (stage d) for the passes, in candidate order
    int np = 0;
    for (int b0 = 0; b0 < nc; b0 += 64) {
      const int c = b0 + lane;
      int p = 0;
      if (c < nc) p = cpass[c] && !rovf[crs[c] >> 1];
      const int incl = wave_scan_incl_i32(p);
      if (c < nc) {
        const int k = p ? np + incl - 1 : -1;
        crec[c] = (int16_t)k;
        if (k >= 0 && k < F_PCAP) rec2cand[k] = (int16_t)c;
      }
      np += __shfl(incl, 63, 64);
    }
    __syncthreads();
    if (np > F_PCAP) {   // too many passes for LDS: every read of the task that has one goes staged
      for (int c = lane; c < nc; c += 64)
        if (crec[c] >= F_PCAP) rovf[crs[c] >> 1] = 1;
      __syncthreads();
    }
    st_cand += (lane == 0) ? (unsigned long long)nc : 0ull;
    st_pass += (lane == 0) ? (unsigned long long)np : 0ull;
    {
      const int npl = np < F_PCAP ? np : F_PCAP;
      // lane l computes records l, l + 64, ...
      for (int k = lane; k < npl; k += 64) {
        const int c = rec2cand[k];
        if (rovf[crs[c] >> 1]) continue;
        const int64_t rs = 2 * base + crs[c];
        recs[k] = dp_one<IM, COLOR>(a.g, a.r, a.sc, (int)rs, cdiag[c], scratch, tbbuf + (int64_t)k * TBB);
      }
    }
    __syncthreads();

    // 4. hit map replay (run.c:660-697, hit_map.c:50-144): one lane per read
    for (int lr = lane; lr < nr; lr += 64) {
      const int64_t read = base + lr;
      if (rovf[lr]) { a.overflow[read] = 1; continue; }
      a.overflow[read] = 0;
      st_hits += (unsigned long long)(rawh[2 * lr] + rawh[2 * lr + 1]);
      Slot s[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) s[k] = a.slots[read * 4 + k];
      const int L = a.r.len[read];
      bool dirty = false;
      for (int strand = 0; strand < 2; ++strand) {
        if (!((a.sense_mask >> strand) & 1)) continue;
        const int j = 2 * lr + strand;
        int T = a.sc.t;
        for (int c = cstart[j]; c < cstart[j + 1]; ++c) {
          const int kr = crec[c];
          if (kr < 0) continue;
          const AlignRec x = recs[kr];
          if (!(x.S >= T && x.score > T)) continue;
          st_acc++;
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
              for (int jj = k; jj > i + 1; --jj) s[jj] = s[jj - 1];
              Slot nw;
              nw.score = score; nw.ref_id = a.ref_id; nw.ref_start = ref_start; nw.ref_end = ref_end;
              nw.read_start = read_start; nw.read_end = read_end; nw.sense = strand ? -1 : 1;
              nw.pass = PASS_FUSED; nw.rec = cell; nw.tb_len = x.tb_len; nw.cell = cell;
              s[i + 1] = nw;
              dirty = true;
              // the traceback goes into the freed cell
              const uint32_t* src = (const uint32_t*)(tbbuf + (int64_t)kr * TBB);
              uint32_t* dst = (uint32_t*)(a.cells + (read * 4 + cell) * (int64_t)TBB);
              for (int w = 0; w < TBB / 4; ++w) dst[w] = src[w];
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
    __syncthreads();
  }
  if (a.stats) {
    if (lane == 0) { atomicAdd(&a.stats[0], st_cand); atomicAdd(&a.stats[1], st_pass); }
    atomicAdd(&a.stats[2], st_acc);
    atomicAdd(&a.stats[3], st_hits);
  }
}

// clears the task counter and the stats inside a captured graph (no memset nodes)
__global__ void k_fused_reset(int32_t* counter, unsigned long long* stats) {
  if (threadIdx.x == 0) { *counter = 0; }
  if (stats && threadIdx.x < 4) stats[threadIdx.x] = 0;
}

hipError_t launch_fused_reset(int32_t* counter, unsigned long long* stats, hipStream_t st) {
  hipLaunchKernelGGL(k_fused_reset, dim3(1), dim3(64), 0, st, counter, stats);
  return hipGetLastError();
}

template <int W, int IM, bool COLOR>
static hipError_t launch_fused_t(const FusedArgs& a, int grid, hipStream_t st) {
  hipLaunchKernelGGL((k_fused<W, IM, COLOR>), dim3(grid), dim3(64), 0, st, a);
  return hipGetLastError();
}

// band classes compiled: W from min(I,15) (run.c:109-127), IM the DP register bound
hipError_t launch_fused(const FusedArgs& a, int grid, hipStream_t st) {
  const int I = a.sc.I, W = a.sc.W;
  const bool c = a.sc.color != 0;
#define STORM_FUSED_CASE(WV, IMV)                                                   \
  if (W == WV && I <= IMV)                                                          \
    return c ? launch_fused_t<WV, IMV, true>(a, grid, st) : launch_fused_t<WV, IMV, false>(a, grid, st);
  STORM_FUSED_CASE(2, 1)
  STORM_FUSED_CASE(4, 3)
  STORM_FUSED_CASE(8, 7)
  STORM_FUSED_CASE(16, 15)
  STORM_FUSED_CASE(16, 31)
#undef STORM_FUSED_CASE
  return hipErrorInvalidValue;
}

// occupancy-sized persistent grid for the band class
int fused_blocks_per_cu(int W, int I, int color) {
  int nb = 0;
  hipError_t e = hipErrorInvalidValue;
#define STORM_OCC(WV, IMV)                                                                                     \
  if (W == WV && I <= IMV) {                                                                                   \
    e = color ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_fused<WV, IMV, true>, 64, 0)               \
              : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_fused<WV, IMV, false>, 64, 0);             \
    return e == hipSuccess ? nb : 0;                                                                           \
  }
  STORM_OCC(2, 1)
  STORM_OCC(4, 3)
  STORM_OCC(8, 7)
  STORM_OCC(16, 15)
  STORM_OCC(16, 31)
#undef STORM_OCC
  return 0;
}

}  // namespace storm
