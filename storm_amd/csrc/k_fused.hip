// k_fused.hip -- persistent fused seed -> lookup -> filter -> DP -> hit map kernel (batched path).
//
// Same semantics as the staged kernels (k_seed, k_filter, k_align, k_replay; run.c:555-747,
// hit_map.c:50-144), in one launch.  Workgroups are single wavefronts that pull tasks from a
// device counter until the batch is exhausted (every wave reaches the exit).  A task is a group
// of G consecutive reads (both strands).  Their candidates accumulate in the workgroup's own
// slice of a global work buffer (diagonal, read strand, filter flag, record index), which stays
// in L2; when the next read might not fit, the accumulated candidates are flushed: filtered two
// per lane, the passes scored one per lane without a direction matrix (stopping as soon as a
// record can no longer be accepted), the few records that can still be accepted aligned again
// with directions and traceback, and the hit maps of the completed reads replayed one lane per
// read over those records only.  So a group of reads keeps all 64 lanes busy when one read strand has only a
// few candidates, and a read with thousands of candidates (a large genome) fills the lanes by
// itself.  Reads of any length share a task; one launch serves one band class (W, I).
// A read whose raw hits exceed the LDS sort buffer, or whose flush has more passes than the
// record buffer, is flagged in `overflow` and left untouched; the host sends it through the
// staged path.
#include "dev_align.h"
#include "dev_align.h"
#include "dev_filter.h"
#include "dev_seed.h"
#include "storm_kernels.h"

// Measurement build (make phases, -DSTORM_FUSED_PHASES): each wavefront sums the constant-clock
// time it spends in the kernel's phases and adds it to FusedArgs::phase at exit; the product build
// compiles none of it.
#ifdef STORM_FUSED_PHASES
#define PH_START() uint64_t ph_t = wall_clock64(); uint64_t ph[FUSED_PHASES] = {}
#define PH_MARK(k) do { const uint64_t t_ = wall_clock64(); ph[k] += t_ - ph_t; ph_t = t_; } while (0)
#define PH_FLUSH() do { if (a.phase && lane == 0) for (int q_ = 0; q_ < FUSED_PHASES; ++q_) atomicAdd(&a.phase[q_], (unsigned long long)ph[q_]); } while (0)
#else
#define PH_START() do {} while (0)
#define PH_MARK(k) do {} while (0)
#define PH_FLUSH() do {} while (0)
#endif

namespace storm {

constexpr int F_HCAP = 2048;         // raw hits of one read strand (LDS sort buffer)
constexpr int F_CCAP = FUSED_CCAP;   // candidates per flush (work buffer)
constexpr int F_PCAP = FUSED_PCAP;   // filter passes (DP records) per flush
constexpr int F_GMAX = 64;           // reads per task
constexpr int F_PENDING = -3;        // Slot.pass while the entry's traceback is still in the flush buffer
static_assert(TBB % 16 == 0, "traceback records are copied in 16-byte units");
static_assert(F_CCAP >= 2 * F_HCAP, "a flush must hold both strands of one read");
static_assert((F_HCAP - F_PCAP) * 4 >= COLOR_TAB, "the color table lives in hbuf past the record lists");

// Wavefronts per SIMD: four for bands up to 15 columns in nucleotide mode (at W = 8 the kernel then
// keeps 128 VGPRs and spills a few bytes, measured 4 % faster on 100-nt reads than three) and for
// color mode at W <= 4 (128 VGPRs; 10 KB of LDS per workgroup keeps 16 workgroups on a CU), three
// for color mode at W = 8 (its color state costs more registers: four measured 2 % slower), two for
// the wide bands (up to 256 VGPRs).
template <int W, int IM, bool COLOR>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W <= 4 ? 4 : W <= 8 ? (COLOR ? 3 : 4) : 2)))
void k_fused(FusedArgs a) {
  __shared__ __align__(16) int32_t hbuf[F_HCAP];
  __shared__ int32_t lk[192];
  __shared__ int32_t cstart[2 * F_GMAX + 1];
  __shared__ int32_t rawh[2 * F_GMAX];
  __shared__ uint8_t rovf[F_GMAX];
  __shared__ int32_t task_sh;
  __shared__ unsigned long long stl[4];   // candidates, filter passes, accepted, raw hits

  const int lane = threadIdx.x;
  const int n = a.n_reads[0];          // the launch maps reads [row0, n) of the batch
  const int64_t row0 = a.n_reads[1];
  const int G = a.group;
  const int I = a.sc.I;
  const FilterConst fk = filter_const(a.sc.fmatch, a.sc.fmismatch, a.sc.fopen, a.sc.fext, a.sc.fthr);
  // direction words of this lane, interleaved with the other lanes' (64 lanes x scratch_stride bytes
  // per workgroup)
  uint32_t* tbw = (uint32_t*)(a.scratch + (int64_t)blockIdx.x * 64 * a.scratch_stride) + lane;
  uint8_t* tbbuf = a.tbbuf + (int64_t)blockIdx.x * F_PCAP * TBB;
  uint8_t* wb = a.work + (int64_t)blockIdx.x * FUSED_WORK_BYTES;
  int32_t* cdiag = (int32_t*)wb;
  int32_t* crec = cdiag + F_CCAP;
  AlignRec* recs = (AlignRec*)(crec + F_CCAP);
  int32_t* rec2cand = (int32_t*)(recs + F_PCAP);
  uint8_t* crs = (uint8_t*)(rec2cand + F_PCAP);
  uint8_t* cpass = crs + F_CCAP;
  if (lane < 4) stl[lane] = 0;

  int64_t base = 0;
  PH_START();
  // filter, DP and hit-map replay of the nc candidates of reads [f0, f1) of the task
  auto flush = [&](int f0, int f1, int nc) {
    if (lane == 0) cstart[2 * f1] = nc;
    __syncthreads();
    PH_MARK(0);
    // filter (stage c): two candidates per lane
    for (int b0 = 0; b0 < nc; b0 += 128) {
      const int c0 = b0 + 2 * lane;
      if (c0 < nc) {
        const int c1 = (c0 + 1 < nc) ? c0 + 1 : c0;
        const int64_t rsa = 2 * base + crs[c0], rsb = 2 * base + crs[c1];
        const int La = a.r.len[rsa >> 1], Lb = a.r.len[rsb >> 1];
        int sa, sb;
        filter_pair<W>(a.g, read_planes(a.r, rsa), read_planes(a.r, rsb), La, Lb,
                       (int64_t)cdiag[c0] - a.sc.d + GPAD, (int64_t)cdiag[c1] - a.sc.d + GPAD, fk, sa, sb);
        // the u8 score of a pass (> thr >= 0, so non-zero), 0 for a rejected candidate
        cpass[c0] = (uint8_t)(sa > fk.thr ? sa : 0);
        if (c1 != c0) cpass[c1] = (uint8_t)(sb > fk.thr ? sb : 0);
      }
    }
    __syncthreads();
    PH_MARK(1);
    // record index of every pass, in candidate order
    int np = 0;
    for (int b0 = 0; b0 < nc; b0 += 64) {
      const int c = b0 + lane;
      int p = 0;
      if (c < nc) p = cpass[c] && !rovf[crs[c] >> 1];
      const int incl = wave_scan_incl_i32(p);
      if (c < nc) {
        const int k = p ? np + incl - 1 : -1;
        crec[c] = k;
        if (k >= 0 && k < F_PCAP) rec2cand[k] = c;
      }
      np += __shfl(incl, 63, 64);
    }
    __syncthreads();
    if (np > F_PCAP) {   // too many passes for the record buffer: their reads go staged
      for (int c = lane; c < nc; c += 64)
        if (crec[c] >= F_PCAP) rovf[crs[c] >> 1] = 1;
      __syncthreads();
    }
    if (lane == 0) { stl[0] += (unsigned long long)nc; stl[1] += (unsigned long long)np; }
    // DP (stage d) over the flush's filter passes
    const int npl = np < F_PCAP ? np : F_PCAP;
    // the record runs (0), is skipped (1: its filter score bounds its DP score at or below t, so
    // it is written as never accepted), or belongs to an overflowing read (2)
    auto triage = [&](int k, int& rs, int& h) -> int {
      const int c = rec2cand[k];
      if (rovf[crs[c] >> 1]) return 2;
      rs = (int)(2 * base + crs[c]);
      h = cdiag[c];
      const int fs = cpass[c];
      if (!COLOR && fs <= a.sc.t && fs < 255 && dp_bounded_by_filter(a.r, DPConst(a.sc), rs, h)) {
        AlignRec x;
        x.score = MIN_SCORE;
        x.S = MIN_SCORE;
        x.win = 0;
        x.best_j = x.best_i0 = x.best_j0 = -1;
        x.tb_len = 0;
        recs[k] = x;
        return 1;
      }
      return 0;
    };
    // the records that run, compacted in record order into hbuf (free during a flush), so that
    // skipped and overflowing records leave no lane idle
    int nrun = 0;
    for (int b0 = 0; b0 < npl; b0 += 64) {
      const int k = b0 + lane;
      int rs = 0, h = 0, r = 0;
      if (k < npl) r = triage(k, rs, h) == 0;
      const int incl = wave_scan_incl_i32(r);
      if (r) hbuf[nrun + incl - 1] = k;
      nrun += __shfl(incl, 63, 64);
    }
    // color DP: the diagonal-move table in the upper half of hbuf, which the flush does not use
    // (its record lists hold at most F_PCAP entries); the next task's candidates overwrite it
    uint8_t* ctab = (uint8_t*)(hbuf + F_PCAP);
    if (COLOR) copy_color_table(ctab, lane, 64);
    __syncthreads();
    PH_MARK(2);
    // One pass per record, lane l taking the running records hbuf[l], hbuf[l + 64], ...: the rows'
    // directions are stored as they are computed (the lane's scratch words), a record stops at the
    // first row where it can no longer be accepted, and one that can (score > t, S >= t) is walked
    // back at once (codes to tbbuf), so no record is aligned twice.
    for (int j = lane; j < nrun; j += 64) {
      const int k = hbuf[j];
      const int c = rec2cand[k];
      recs[k] = dp_one<IM, COLOR, true>(a.g, a.r, a.sc, (int)(2 * base + crs[c]), cdiag[c], tbw,
                                        tbbuf + (int64_t)k * TBB, ctab);
    }
    __syncthreads();   // hbuf is rewritten below
    PH_MARK(3);
    // the records that can still be accepted (score > t, S >= t), ascending, as
    // (read strand << 16) | record in the hit buffer
    int nv = 0;
    for (int b0 = 0; b0 < npl; b0 += 64) {
      const int k = b0 + lane;
      int v = 0, j = 0;
      if (k < npl) {
        j = crs[rec2cand[k]];
        if (!rovf[j >> 1]) {
          const AlignRec x = recs[k];
          v = x.score > a.sc.t && x.S >= a.sc.t;
        }
      }
      const int incl = wave_scan_incl_i32(v);
      if (v) hbuf[nv + incl - 1] = (j << 16) | k;
      nv += __shfl(incl, 63, 64);
    }
    __syncthreads();
    // hit map replay (run.c:660-697, hit_map.c:50-144): one lane per read
    int st_acc = 0, st_hits = 0;
    for (int lr = f0 + lane; lr < f1; lr += 64) {
      const int64_t read = base + lr;
      if (rovf[lr]) { a.overflow[read] = 1; continue; }
      a.overflow[read] = 0;
      st_hits += rawh[2 * lr] + rawh[2 * lr + 1];
      Slot s[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) s[k] = a.slots[read * 4 + k];
      const int L = a.r.len[read];
      bool dirty = false;
      // this read's viable records: the entries of strand 2lr, then of strand 2lr + 1
      int e = 0;
      {
        int lo = 0, hi = nv;   // first entry of strand >= 2lr
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if ((hbuf[mid] >> 16) < 2 * lr) lo = mid + 1;
          else hi = mid;
        }
        e = lo;
      }
      for (int strand = 0; strand < 2; ++strand) {
        const int j = 2 * lr + strand;
        int T = a.sc.t;
        for (; e < nv && (hbuf[e] >> 16) == j; ++e) {
          const int kr = hbuf[e] & 0xffff;
          const AlignRec x = recs[kr];
          if (!(x.S >= T && x.score > T)) continue;
          st_acc++;
          const int score = x.score;
          const int ref_start = x.win + x.best_j0, ref_end = x.win + x.best_j;
          const int read_start = x.best_i0, read_end = L - 1;
          // hit_map__update with compile-time slot indices only (the slots stay in registers)
          int k = 4;
          bool found = false, drop = false;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            if (!found) {
              if (s[q].ref_id == -1) {
                k = q;
                found = true;
              } else if (s[q].ref_id == a.ref_id &&
                         ((s[q].ref_start == ref_start && s[q].read_start == read_start) ||
                          (s[q].ref_end == ref_end && s[q].read_end == read_end))) {
                k = q;
                found = true;
                drop = !(score > s[q].score);
              }
            }
          }
          if (!drop) {
            if (k == 4) k = 3;
            int i = k;   // i = k; while (i >= 0 && score > s[i].score) --i;
            bool go = true;
#pragma unroll
            for (int q = 3; q >= 0; --q) {
              if (go && q <= k) {
                if (score > s[q].score) i = q - 1;
                else go = false;
              }
            }
            if (i != 3) {
              const int p = i + 1;
              int32_t cell = 0;
#pragma unroll
              for (int q = 0; q < 4; ++q) if (q == k) cell = s[q].cell;
#pragma unroll
              for (int q = 3; q >= 1; --q) if (q > p && q <= k) s[q] = s[q - 1];
              Slot nw;
              nw.score = score; nw.ref_id = a.ref_id; nw.ref_start = ref_start; nw.ref_end = ref_end;
              nw.read_start = read_start; nw.read_end = read_end; nw.sense = strand ? -1 : 1;
              // the traceback stays in this flush's buffer (record kr) until the read is done
              nw.pass = F_PENDING; nw.rec = kr; nw.tb_len = x.tb_len; nw.cell = cell;
#pragma unroll
              for (int q = 0; q < 4; ++q) if (q == p) s[q] = nw;
              dirty = true;
            }
          }
          if (s[3].score > T) T = s[3].score;
        }
      }
      if (dirty) {
        // tracebacks of the entries this flush inserted and kept, each into its own cell (the four
        // slots hold a permutation of the read's four cells), then the slots
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          if (s[k].pass == F_PENDING) {
            const uint4* src = (const uint4*)(tbbuf + (int64_t)s[k].rec * TBB);
            uint4* dst = (uint4*)(a.cells + (read * 4 + s[k].cell) * (int64_t)TBB);
            const int n16 = ((s[k].tb_len + 1) / 2 + 15) / 16;
            for (int w = 0; w < n16; ++w) dst[w] = src[w];
            s[k].pass = PASS_CELL;
            s[k].rec = s[k].cell;
          }
          a.slots[read * 4 + k] = s[k];
        }
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      st_acc += __shfl_xor(st_acc, o, 64);
      st_hits += __shfl_xor(st_hits, o, 64);
    }
    if (lane == 0) { stl[2] += (unsigned long long)st_acc; stl[3] += (unsigned long long)st_hits; }
    __syncthreads();
    PH_MARK(4);
  };

  for (;;) {
    if (lane == 0) task_sh = atomicAdd(a.task_counter, 1);
    __syncthreads();
    const int task = task_sh;
    __syncthreads();
    base = row0 + (int64_t)task * G;
    PH_MARK(5);
    if (base >= n) break;
    const int nr = (n - base) < G ? (int)(n - base) : G;
    for (int j = lane; j < nr; j += 64) rovf[j] = 0;
    __syncthreads();

    // candidates of every read strand (stages a + b), flushed whenever the next read might not
    // fit and after the last read (one call site keeps the flush inlined once)
    int nc = 0, f0 = 0;
    for (int lr = 0;; ++lr) {
      if (lr == nr || nc + 2 * F_HCAP > F_CCAP) {
        flush(f0, lr, nc);
        nc = 0;
        f0 = lr;
        if (lr == nr) break;
      }
      for (int strand = 0; strand < 2; ++strand) {
        const int j = 2 * lr + strand;
        if (lane == 0) { cstart[j] = nc; rawh[j] = 0; }
        const bool wanted = ((a.sense_mask >> strand) & 1) && !rovf[lr];
        if (!wanted) continue;
        int raw = 0;
        // the register sort of 2048 hits where it keeps the occupancy: nucleotide bands of 8+
        // diagonal pairs (measured 5 % faster on 100 nt reads of a repeat genome) and the wide
        // bands, which run two wavefronts per SIMD anyway (color at W = 4 lost a wavefront: 12 %)
        const int cnt = wave_candidates<(W >= 8 && !COLOR) || W >= 16>(a.r, a.seeds, 2 * base + j, I, hbuf, F_HCAP, lk,
                                                                        cdiag + nc, F_CCAP - nc, &raw);
        if (lane == 0) rawh[j] = raw;
        if (cnt < 0) {
          if (lane == 0) rovf[lr] = 1;
          __syncthreads();
          continue;
        }
        for (int e = lane; e < cnt; e += 64) crs[nc + e] = (uint8_t)j;
        nc += cnt;
      }
    }
  }
  if (a.stats && lane < 4) atomicAdd(&a.stats[lane], stl[lane]);
  PH_FLUSH();
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
