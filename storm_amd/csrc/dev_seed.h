// dev_seed.h -- spaced-seed keys, CSR bucket gathers and the candidate keep rule for one read
// strand held by one wavefront (used by k_fused; k_seed.hip holds the staged variant).  The
// functions synchronise with __syncthreads(): they are called from one-wavefront workgroups.
//
// Semantics: run.c:570-698 (heap merge of the per-(seed, position) hit lists), seed.c:203-214
// (keys), index.c:212-241 (buckets): the candidates of a read strand are the distinct diagonals
// q - rp in ascending order, the smallest one dropped iff it is <= -I.
#pragma once
#include "storm_dev.h"

namespace storm {

// wave-wide inclusive scan (64 lanes)
__device__ __forceinline__ int wave_scan_incl_i32(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int n = __shfl_up(v, o, 64);
    if (lane >= o) v += n;
  }
  return v;
}

// bitonic sort of buf[0..n) in place (one wavefront, buf in LDS, padded to a power of two)
__device__ __forceinline__ void wave_bitonic_sort(int32_t* buf, int n) {
  const int lane = threadIdx.x & 63;
  int np2 = 1;
  while (np2 < n) np2 <<= 1;
  for (int e = n + lane; e < np2; e += 64) buf[e] = 0x7fffffff;
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
}

// Sorts buf[0..n) (LDS, n <= 64 E) in registers and applies the keep rule: E elements per lane,
// element lane*E + k in v[k].  Bitonic network over 64 E keys padded with INT_MAX: compare-exchange
// stages at distance j < E stay inside a lane; longer ones pair lane with lane ^ (j / E) through a
// cross-lane shuffle, so no stage waits on LDS or a barrier.  The kept diagonals (first one only if
// > -I, then each one larger than its predecessor) go to out[0..) in order; returns their number,
// or -1 when more than cap.
template <int E>
__device__ int wave_sort_keep(const int32_t* buf, int n, int I, int32_t* out, int cap) {
  const int lane = threadIdx.x & 63;
  int32_t v[E];
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int e = lane * E + k;
    v[k] = e < n ? buf[e] : 0x7fffffff;
  }
#pragma unroll
  for (int kk = 2; kk <= 64 * E; kk <<= 1) {
#pragma unroll
    for (int j = kk >> 1; j > 0; j >>= 1) {
      if (j >= E) {
        const int lj = j / E;
        const bool lower = (lane & lj) == 0;
#pragma unroll
        for (int k = 0; k < E; ++k) {
          const bool up = ((lane * E + k) & kk) == 0;
          const int32_t w = __shfl_xor(v[k], lj, 64);
          v[k] = (lower == up) ? min(v[k], w) : max(v[k], w);
        }
      } else {
#pragma unroll
        for (int k = 0; k < E; ++k) {
          if ((k & j) == 0) {
            const bool up = ((lane * E + k) & kk) == 0;
            const int32_t a = v[k], b = v[k + j];
            v[k] = up ? min(a, b) : max(a, b);
            v[k + j] = up ? max(a, b) : min(a, b);
          }
        }
      }
    }
  }
  const int32_t prev_last = __shfl_up(v[E - 1], 1, 64);
  int cnt = 0;
  uint32_t keep = 0;
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int e = lane * E + k;
    const int32_t p = k ? v[k - 1] : prev_last;
    const bool kp = (e < n) && ((e == 0) ? (v[k] > -I) : (v[k] > p));
    keep |= (uint32_t)kp << k;
    cnt += kp;
  }
  const int incl = wave_scan_incl_i32(cnt);
  const int total = __shfl(incl, 63, 64);
  if (total > cap) return -1;
  int o = incl - cnt;
#pragma unroll
  for (int k = 0; k < E; ++k)
    if ((keep >> k) & 1u) out[o++] = v[k];
  return total;
}

// All candidates of read strand rs, ascending, into out[0..), one wavefront.  hbuf is an LDS
// buffer of `hcap` entries (hcap >= 64, power of two), lk an LDS buffer of 192 ints.  The raw hits
// of 64 lookups are gathered cooperatively: hit e of the batch goes to lane e mod 64, which finds
// its bucket by binary search over the lookups' inclusive counts, so consecutive lanes read
// consecutive entries of a bucket.  Returns the number of candidates, or -1 when the raw hits
// exceed hcap or the candidates exceed `cap` (the caller sends the read to the staged path, which
// has no such limits).  *raw_hits receives the bucket entries read (the lookups' total count).
// REG2048: sort up to 2048 raw hits in registers (32 per lane) rather than in LDS, five times
// fewer instructions for reads with many hits (repeats), but it raises the caller's register
// count; the callers whose occupancy it would lower leave it off.
template <bool REG2048>
__device__ int wave_candidates(const ReadsDev& r, const SeedSetDev& seeds, int64_t rs, int I, int32_t* hbuf, int hcap,
                               int32_t* lk, int32_t* out, int cap, int* raw_hits) {
  const int lane = threadIdx.x & 63;
  const int L = r.len[rs >> 1];
  const uint32_t* pl = read_planes(r, rs);
  // lookups of one batch (seed s, positions k0 .. k0+63): position and the bucket bounds
  // first[key], first[key+1]; issued one batch ahead, so their latency overlaps the batch before
  auto lookup = [&](int s, int k0, int& rp, int& b0, int& b1) {
    const SeedDev& sd = seeds.s[s];
    const int k = k0 + lane;
    rp = 0;
    b0 = b1 = 0;
    if (k < seed_npos(sd, L)) {
      rp = sd.n_pos > 0 ? sd.pos[k] : k;
      const int key = seed_key(sd, pl, rp);
      b0 = sd.first[key];
      b1 = sd.first[key + 1];
    }
  };
  auto next_batch = [&](int& s, int& k0) {
    k0 += 64;
    while (s < seeds.n && k0 >= seed_npos(seeds.s[s], L)) { ++s; k0 = 0; }
  };
  int run = 0;
  bool over = false;
  int s = 0, k0 = -64;
  next_batch(s, k0);
  int rp = 0, b0 = 0, b1 = 0;
  if (s < seeds.n) lookup(s, k0, rp, b0, b1);
  while (s < seeds.n) {
    int s2 = s, k2 = k0;
    next_batch(s2, k2);
    int rp2 = 0, b02 = 0, b12 = 0;
    if (s2 < seeds.n) lookup(s2, k2, rp2, b02, b12);
    const SeedDev& sd = seeds.s[s];
    const int c = b1 - b0;
    const int incl = wave_scan_incl_i32(c);
    const int tot = __shfl(incl, 63, 64);
    if (!over && run + tot <= hcap) {
      lk[lane] = b0 - (incl - c);   // bucket entry of batch hit e is lk[q] + e
      lk[64 + lane] = rp;
      lk[128 + lane] = incl;
      __syncthreads();
      // hit e of the batch: lookup q = the first with incl[q] > e; four hits per lane at a time,
      // so four bucket reads are in flight together
      for (int e0 = lane; e0 < tot; e0 += 256) {
        int q[4], v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int e = e0 + 64 * u;
          int qq = 0;
#pragma unroll
          for (int step = 32; step > 0; step >>= 1)
            if (lk[128 + qq + step - 1] <= e) qq += step;
          q[u] = qq;
          v[u] = e < tot ? sd.map[lk[qq] + e] : 0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int e = e0 + 64 * u;
          if (e < tot) hbuf[run + e] = v[u] - lk[64 + q[u]];
        }
      }
      __syncthreads();
    } else {
      over = true;
    }
    run += tot;
    s = s2;
    k0 = k2;
    rp = rp2;
    b0 = b02;
    b1 = b12;
  }
  *raw_hits = run;
  if (over) return -1;
  if (run == 0) return 0;
  if (run <= 64) return wave_sort_keep<1>(hbuf, run, I, out, cap);
  if (run <= 128) return wave_sort_keep<2>(hbuf, run, I, out, cap);
  if (run <= 256) return wave_sort_keep<4>(hbuf, run, I, out, cap);
  if (run <= 512) return wave_sort_keep<8>(hbuf, run, I, out, cap);
  if (run <= 1024) return wave_sort_keep<16>(hbuf, run, I, out, cap);
  if (REG2048 && run <= 2048) return wave_sort_keep<32>(hbuf, run, I, out, cap);
  wave_bitonic_sort(hbuf, run);
  int placed = 0;
  for (int e0 = 0; e0 < run; e0 += 64) {
    const int e = e0 + lane;
    int keep = 0;
    int32_t v = 0;
    if (e < run) {
      v = hbuf[e];
      keep = (e == 0) ? (v > -I) : (v > hbuf[e - 1]);
    }
    const int incl = wave_scan_incl_i32(keep);
    if (keep && placed + incl - 1 < cap) out[placed + incl - 1] = v;
    placed += __shfl(incl, 63, 64);
  }
  __syncthreads();
  return placed > cap ? -1 : placed;
}

}  // namespace storm
