// capi.cpp -- C ABI of the gfx950 mapping path (include/storm_gpu.h): device buffers, the
// per-batch pipeline (seed -> lookup -> candidates -> filter -> DP -> hit map) and the drop-in
// rows of the reference's filter table.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <functional>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "host_pool.h"
#include "mapper.h"

namespace storm {

static thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ------------------------------------------------------------------------------ genome planes

int filter_lanes(int d) { return d <= 1 ? 2 : d <= 3 ? 4 : d <= 7 ? 8 : 16; }   // run.c:109-127

int tb_ranks(const storm_params& p) { return p.max_hits == 0 ? 4 : p.max_hits; }

void fill_score(const storm_params& p, ScoreDev& sc) {
  sc.fmatch = std::abs(p.match) & 0xff;
  sc.fmismatch = std::abs(p.mismatch) & 0xff;
  sc.fopen = std::abs(p.gap_open) & 0xff;
  sc.fext = std::abs(p.gap_extend) & 0xff;
  unsigned u = (unsigned)p.z;
  if (u >= 255) u = 254;
  sc.fthr = (int)u;
  for (int q = 0; q < 4; ++q) {   // SCALE_SCORE (alignment.h:30), alignment.c:276-279
    const short m = (short)p.match, x = (short)p.mismatch;
    sc.ps[0][q] = (int)((m * (q + 1) + 0.) / 4);
    sc.ps[1][q] = (int)((x * (q + 1) + 0.) / 4);
  }
  sc.gap_open = (short)p.gap_open;
  sc.gap_ext = (short)p.gap_extend;
  sc.color = p.color;
  sc.I = p.allowed_indels;
  sc.d = p.allowed_indels >= 16 ? 15 : p.allowed_indels;
  sc.W = filter_lanes(sc.d);
  sc.t = p.t;
}


int ensure_ref_resident(storm_mapper* m, size_t r) {
  RefState& rs = *m->res->rstate[r];
  if (rs.resident) return STORM_OK;
  const HostRef& ref = m->res->refs[r];
  const double t0 = now_s();
  DevMem dcode;
  int rc = STORM_OK;
  if (!rs.g.lh.p) {   // planes and the codes the index build reads
    rc = upload_genome(ref.code.data(), ref.mask.empty() ? nullptr : ref.mask.data(), ref.size, rs.g, m->st, &dcode);
    if (rc) return rc;
  } else {            // planes kept from an earlier pass: the codes only
    HIPOK(dcode.ensure((size_t)ref.size + 64));
    if (ref.size > 0) HIPOK(hipMemcpyAsync(dcode.p, ref.code.data(), (size_t)ref.size, hipMemcpyHostToDevice, m->st));
  }
  rc = build_indexes(ref, dcode, m->res->seeds, m->p.stdev_distance, rs.idx, m->st);
  if (rc) return rc;
  m->res->index_seconds += now_s() - t0;
  rs.resident = true;
  return STORM_OK;
}

// the seed indexes of a sequence are released after its pass; its planes (0.375 byte per code) stay
// for the SAM output and the next call
static void release_ref(storm_mapper* m, size_t r) {
  RefState& rs = *m->res->rstate[r];
  rs.idx.clear();
  rs.resident = false;
}

int ensure_planes(storm_mapper* m) {
  for (size_t r = 0; r < m->res->refs.size(); ++r) {
    RefState& rs = *m->res->rstate[r];
    if (rs.g.lh.p) continue;
    const HostRef& ref = m->res->refs[r];
    const int rc = upload_genome(ref.code.data(), ref.mask.empty() ? nullptr : ref.mask.data(), ref.size, rs.g, m->st);
    if (rc) return rc;
  }
  return STORM_OK;
}

static SeedSetDev seed_set(const RefState& rs) {
  SeedSetDev s{};
  s.n = (int)rs.idx.size();
  for (int k = 0; k < s.n; ++k) {
    const SeedSpec& sp = rs.idx[k].spec;
    s.s[k].weight = sp.weight;
    s.s[k].span = sp.span;
    s.s[k].n_pos = (int)sp.pos.size();
    for (int j = 0; j < sp.weight; ++j) s.s[k].offs[j] = sp.offs[j];
    for (size_t j = 0; j < sp.pos.size(); ++j) s.s[k].pos[j] = sp.pos[j];
    s.s[k].first = rs.idx[k].first.as<int32_t>();
    s.s[k].map = rs.idx[k].map.as<int32_t>();
  }
  return s;
}

// Stable order of the view's reads by length (counting sort: lengths are 1..MAXL); identity
// reports whether that order is the view's own.
// Longest reads first (stable within a length): the persistent kernel hands out the costliest
// tasks first and the short ones fill the tail, and equal lengths stay contiguous for the plain path.
// Device order: the reads of each chunk [bounds[k], bounds[k + 1]) of the view by length within
// the chunk, longest first, stable (a counting sort), so that chunk k's device rows are its own
// reads; order32, when given, receives the same order in 32 bits (for the device).  hist counts the
// reads per length over the whole view (lengths outside 0 .. MAXL + 1 at its ends), identity tells
// whether the order is the view's own.  Large views are counted and scattered by the host pool,
// each thread a contiguous part of the chunk (the part's reads of one length keep their order, and
// the parts are placed in turn within each length).
static void length_order_chunks(const storm_reads_view* v, const std::vector<int64_t>& bounds, std::vector<int64_t>& order,
                                bool& identity, std::vector<int64_t>& hist, std::vector<int32_t>* order32 = nullptr) {
  const int64_t n = v->n_reads;
  constexpr int NB = MAXL + 2;   // bucket b holds length MAXL + 1 - b
  auto bucket = [&](int64_t i) { return (int)(MAXL + 1 - std::min(std::max((int)v->len[i], 0), MAXL + 1)); };
  order.resize((size_t)n);
  if (order32) order32->resize((size_t)n);
  hist.assign(NB, 0);
  identity = true;
  for (size_t k = 0; k + 1 < bounds.size(); ++k) {
    const int64_t lo = bounds[k], hi = bounds[k + 1];
    const int T = (int)std::min<int64_t>(8, std::max<int64_t>(1, (hi - lo) >> 16));
    std::vector<int64_t> cnt((size_t)T * NB, 0);
    std::vector<uint8_t> sorted((size_t)T, 1);
    auto part = [&](int t, int64_t& a, int64_t& b) {
      a = lo + (hi - lo) * t / T;
      b = lo + (hi - lo) * (t + 1) / T;
    };
    auto count = [&](int t) {
      int64_t a, b;
      part(t, a, b);
      int64_t* c = cnt.data() + (size_t)t * NB;
      for (int64_t i = a; i < b; ++i) {
        c[bucket(i)]++;
        if (i > a && v->len[i] > v->len[i - 1]) sorted[(size_t)t] = 0;
      }
    };
    if (T > 1) HostPool::get().run(T, count);
    else count(0);
    for (int t = 0; t < T; ++t) {
      int64_t a, b;
      part(t, a, b);
      if (!sorted[(size_t)t] || (t > 0 && a < b && a > lo && v->len[a] > v->len[a - 1])) identity = false;
    }
    // offsets, bucket-major then part: off[t][b] = lo + reads in buckets < b + reads of bucket b
    // in parts < t
    int64_t run = lo;
    for (int b = 0; b < NB; ++b)
      for (int t = 0; t < T; ++t) {
        const int64_t c = cnt[(size_t)t * NB + b];
        cnt[(size_t)t * NB + b] = run;
        run += c;
        hist[(size_t)(MAXL + 1 - b)] += c;
      }
    auto scatter = [&](int t) {
      int64_t a, b;
      part(t, a, b);
      int64_t* off = cnt.data() + (size_t)t * NB;
      if (order32)
        for (int64_t i = a; i < b; ++i) {
          const int64_t o = off[bucket(i)]++;
          order[(size_t)o] = i;
          (*order32)[(size_t)o] = (int32_t)i;
        }
      else
        for (int64_t i = a; i < b; ++i) order[(size_t)off[bucket(i)]++] = i;
    };
    if (T > 1) HostPool::get().run(T, scatter);
    else scatter(0);
  }
}


// Uploads the view's reads and arranges them on the device in order[0..) (a permutation of the
// view).  The caller's arrays go to the device as they are; a reordering is a device gather.
// memcpy in parts of 1 MiB or more on up to 8 host threads (a copy between pinned memory and the
// caller's pageable memory runs at the host's memory bandwidth only with several threads)
static void host_copy(void* dst, const void* src, size_t bytes) {
  const int nt = (int)std::min<size_t>(8, std::max<size_t>(1, bytes >> 20));
  if (nt <= 1) {
    memcpy(dst, src, bytes);
    return;
  }
  const size_t per = (bytes + nt - 1) / nt;
  HostPool::get().run(nt, [&](int t) {
    const size_t off = (size_t)t * per;
    if (off < bytes) memcpy((char*)dst + off, (const char*)src + off, std::min(per, bytes - off));
  });
}

// A copy of `bytes` from src to dst; with unit > 0 (device to host only) the source is records of
// `unit` bytes that land dst_unit >= unit bytes apart in dst, the rest of each zero-filled.
struct CopySeg {
  void* dst;
  const void* src;
  size_t bytes;
  size_t unit = 0, dst_unit = 0;
};

// rows records of `unit` bytes from src into dst_unit-byte slots of dst (the rest of each slot
// zeroed), on up to 8 host threads
static void host_expand(void* dst, const void* src, size_t rows, size_t unit, size_t dst_unit) {
  const int nt = (int)std::min<size_t>(8, std::max<size_t>(1, (rows * dst_unit) >> 20));
  auto part = [&](int t) {
    const size_t a = rows * t / nt, b = rows * (t + 1) / nt;
    for (size_t r = a; r < b; ++r) {
      char* d = (char*)dst + r * dst_unit;
      memcpy(d, (const char*)src + r * unit, unit);
      memset(d + unit, 0, dst_unit - unit);
    }
  };
  if (nt <= 1) part(0);
  else HostPool::get().run(nt, part);
}

// Copies between the caller's (pageable) host memory and device memory on stream st (default: the
// mapper's) at pinned-copy speed: the segments are cut into pieces of up to 8 MiB that pass through two pinned
// halves in turn, and the host threads copying one piece overlap the DMA of the next.
// to_host: device -> host, returns when every dst holds its bytes; else host -> device, returns
// once every byte is in pinned memory or on the device (the DMA completes in stream order; the
// halves are not reused before their events say so).
static int staged_copy(storm_mapper* m, const std::vector<CopySeg>& segs, bool to_host, hipStream_t st = nullptr) {
  if (!st) st = m->st;
  const size_t CH = (size_t)8 << 20;
  HIPOK(m->out_stage.ensure(2 * CH));
  for (int k = 0; k < 2; ++k) {
    if (!m->stage_ev[k]) HIPOK(hipEventCreateWithFlags(&m->stage_ev[k], hipEventDisableTiming));
    HIPOK(hipStreamWaitEvent(st, m->stage_ev[k], 0));   // the half's last copy, on whichever stream
  }
  struct Piece {
    char* dst;
    const char* src;
    size_t len;
    size_t unit, dst_unit;
  };
  std::vector<Piece> ps;
  for (const CopySeg& g : segs) {
    if (g.unit && (!to_host || g.dst_unit < g.unit || g.unit > CH || g.bytes % g.unit))
      return fail(STORM_ERR_INVALID, "bad record copy");
    const size_t step = g.unit ? CH / g.unit * g.unit : CH;   // pieces end on record boundaries
    for (size_t off = 0; off < g.bytes; off += step) {
      const size_t doff = g.unit ? off / g.unit * g.dst_unit : off;
      ps.push_back({(char*)g.dst + doff, (const char*)g.src + off, std::min(step, g.bytes - off), g.unit, g.dst_unit});
    }
  }
  for (size_t c = 0; c <= ps.size(); ++c) {
    if (c < ps.size()) {
      char* half = (char*)m->out_stage.p + (c & 1) * CH;
      if (to_host) {   // piece c into half c & 1 (its previous piece, c - 2, was copied out already)
        HIPOK(hipMemcpyAsync(half, ps[c].src, ps[c].len, hipMemcpyDeviceToHost, st));
      } else {         // the half's previous DMA has read it; fill it and send it
        HIPOK(hipEventSynchronize(m->stage_ev[c & 1]));
        host_copy(half, ps[c].src, ps[c].len);
        HIPOK(hipMemcpyAsync(ps[c].dst, half, ps[c].len, hipMemcpyHostToDevice, st));
      }
      HIPOK(hipEventRecord(m->stage_ev[c & 1], st));
    }
    if (to_host && c >= 1) {   // piece c - 1 out of its half while piece c is in flight
      const size_t q = c - 1;
      HIPOK(hipEventSynchronize(m->stage_ev[q & 1]));
      const char* half = (const char*)m->out_stage.p + (q & 1) * CH;
      if (ps[q].unit) host_expand(ps[q].dst, half, ps[q].len / ps[q].unit, ps[q].unit, ps[q].dst_unit);
      else host_copy(ps[q].dst, half, ps[q].len);
    }
  }
  return STORM_OK;
}

int upload_reads(storm_mapper* m, const storm_reads_view* v, const std::vector<int64_t>& order, DevReads& d,
                 const std::vector<int64_t>* hist, const bool* identity_in, const std::vector<int32_t>* order32) {
  m->order_dev_n = -1;
  const int64_t n = v->n_reads;
  if ((int64_t)order.size() != n) return fail(STORM_ERR_INVALID, "read order does not cover the view");
  const int w = filter_lanes(m->p.allowed_indels >= 16 ? 15 : m->p.allowed_indels);
  // lengths outside [w, MAXL] are rejected; with the call's length histogram the reads are only
  // scanned (for the message) when one is
  bool bad = !hist;
  if (hist)
    for (int L = 0; L <= MAXL + 1; ++L)
      if ((*hist)[(size_t)L] && (L <= 0 || L > MAXL || L < w)) bad = true;
  bool identity = identity_in ? *identity_in : true;
  if (bad || !identity_in)
    for (int64_t i = 0; i < n; ++i) {
      const int L = v->len[i];
      if (L <= 0 || L > MAXL)
        return fail(STORM_ERR_INVALID, "read %lld has length %d (1..%d supported)", (long long)i, L, MAXL);
      if (L < w)
        return fail(STORM_ERR_INVALID, "read %lld is shorter (%d) than the filter band lanes (%d)", (long long)i, L, w);
      if (order[(size_t)i] != i) identity = false;
    }
  d.n = n;
  d.has_qual = v->qual != nullptr;
  const int64_t cs = v->code_stride, qs = v->qual ? v->qual_stride : 0;
  const bool has_fq = v->first_qual != nullptr;
  HIPOK(d.codes.ensure((size_t)(n * cs)));
  HIPOK(d.len.ensure((size_t)n * 4));
  if (qs) HIPOK(d.qual.ensure((size_t)(n * qs)));
  if (has_fq) HIPOK(d.first_qual.ensure((size_t)n));
  DevReads& src = identity ? d : m->stage;
  if (!identity) {
    HIPOK(src.codes.ensure((size_t)(n * cs)));
    HIPOK(src.len.ensure((size_t)n * 4));
    if (qs) HIPOK(src.qual.ensure((size_t)(n * qs)));
    if (has_fq) HIPOK(src.first_qual.ensure((size_t)n));
  }
  std::vector<CopySeg> up = {{src.codes.p, v->codes, (size_t)(n * cs)}, {src.len.p, v->len, (size_t)n * 4}};
  if (qs) up.push_back({src.qual.p, v->qual, (size_t)(n * qs)});
  if (has_fq) up.push_back({src.first_qual.p, v->first_qual, (size_t)n});
  int rc = staged_copy(m, up, false);
  if (rc) return rc;
  if (!identity) {
    std::vector<int32_t> o32;
    const int32_t* oh = order32 && (int64_t)order32->size() == n ? order32->data() : nullptr;
    if (!oh) {
      o32.resize((size_t)n);
      for (int64_t i = 0; i < n; ++i) o32[(size_t)i] = (int32_t)order[(size_t)i];
      oh = o32.data();
    }
    HIPOK(m->order_dev.ensure((size_t)n * 4));
    HIPOK(hipMemcpyAsync(m->order_dev.p, oh, (size_t)n * 4, hipMemcpyHostToDevice, m->st));
    m->order_dev_n = n;
    const int32_t* od = m->order_dev.as<int32_t>();
    HIPOK(launch_permute_bytes(src.codes.as<uint8_t>(), cs, od, n, d.codes.as<uint8_t>(), m->st));
    HIPOK(launch_permute_i32(src.len.as<int32_t>(), od, n, d.len.as<int32_t>(), m->st));
    if (qs) HIPOK(launch_permute_bytes(src.qual.as<uint8_t>(), qs, od, n, d.qual.as<uint8_t>(), m->st));
    if (has_fq) HIPOK(launch_permute_bytes(src.first_qual.as<uint8_t>(), 1, od, n, d.first_qual.as<uint8_t>(), m->st));
    HIPOK(hipStreamSynchronize(m->st));   // o32 is released on return
  }
  return prep_device_reads(m, n, cs, qs, qs != 0, has_fq, d);
}

// the slots (device order `order`) into the caller's four arrays (view order); returns mapped reads
static int hits_to_host(storm_mapper* m, const Slot* slots, const std::vector<int64_t>& order, int64_t n,
                        const HitsOut& ho, int64_t& mapped) {
  bool identity = true;
  for (int64_t i = 0; i < n && identity; ++i) identity = order[(size_t)i] == i;
  if (!identity && m->order_dev_n != n) {   // the upload left no order on the device
    std::vector<int32_t> o32((size_t)n);
    for (int64_t i = 0; i < n; ++i) o32[(size_t)i] = (int32_t)order[(size_t)i];
    HIPOK(m->order_dev.ensure((size_t)n * 4));
    HIPOK(hipMemcpyAsync(m->order_dev.p, o32.data(), (size_t)n * 4, hipMemcpyHostToDevice, m->st));
    HIPOK(hipStreamSynchronize(m->st));
    m->order_dev_n = n;
  }
  HIPOK(m->hits_dev.ensure((size_t)(4 * n) * 16));
  int32_t* hd = m->hits_dev.as<int32_t>();
  HIPOK(launch_hits_out(slots, identity ? nullptr : m->order_dev.as<int32_t>(), n, hd, hd + 4 * n, hd + 8 * n,
                        hd + 12 * n, m->st));
  std::vector<int32_t> ref_tmp;
  int32_t* ref_host = ho.ref_id;
  if (!ref_host) { ref_tmp.resize((size_t)(4 * n)); ref_host = ref_tmp.data(); }
  const size_t bytes = (size_t)(4 * n) * 4;
  std::vector<CopySeg> out = {{ref_host, hd + 4 * n, bytes}};
  if (ho.score) out.push_back({ho.score, hd, bytes});
  if (ho.ref_start) out.push_back({ho.ref_start, hd + 8 * n, bytes});
  if (ho.sense) out.push_back({ho.sense, hd + 12 * n, bytes});
  const int rc = staged_copy(m, out, true);
  if (rc) return rc;
  mapped = 0;
  for (int64_t i = 0; i < n; ++i) mapped += ref_host[(size_t)(4 * i)] != -1;
  return STORM_OK;
}

// records per k_align launch: whole wavefronts (the direction words of a wavefront's 64 records
// are interleaved) within a 256 MiB direction scratch
static int64_t align_chunk(int64_t scratch_stride) {
  const int64_t budget = 256LL << 20;
  return std::max<int64_t>(ALIGN_WAVE_RECORDS,
                           (budget / std::max<int64_t>(scratch_stride, 1)) & ~(int64_t)(ALIGN_WAVE_RECORDS - 1));
}

int64_t align_scratch_bytes(int I, int maxL, int64_t n_rec) {
  const int64_t stride = (int64_t)maxL * dp_dir_words(I) * 4;
  const int64_t chunk_alloc = (std::min(align_chunk(stride), std::max<int64_t>(n_rec, 1)) + ALIGN_WAVE_RECORDS - 1) &
                              ~(int64_t)(ALIGN_WAVE_RECORDS - 1);
  return chunk_alloc * stride;
}

int align_records(AlignArgs aa, int maxL, int64_t n_rec, DevMem& scratch, hipStream_t st) {
  aa.scratch_stride = (int64_t)maxL * dp_dir_words(aa.sc.I) * 4;
  const int64_t chunk = align_chunk(aa.scratch_stride);
  HIPOK(scratch.ensure((size_t)align_scratch_bytes(aa.sc.I, maxL, n_rec)));
  aa.scratch = scratch.as<uint8_t>();
  for (int64_t b0 = 0; b0 < n_rec; b0 += chunk) {
    aa.base = b0;
    aa.n = std::min(chunk, n_rec - b0);
    HIPOK(launch_align(aa, st));
  }
  return STORM_OK;
}

int prep_device_reads(storm_mapper* m, int64_t n, int64_t code_stride, int64_t qual_stride, bool has_qual,
                      bool has_first_qual, DevReads& d) {
  d.n = n;
  d.has_qual = has_qual ? 1 : 0;
  HIPOK(d.planes.ensure((size_t)(2 * n) * 2 * RW * 4));
  HIPOK(d.qplanes.ensure((size_t)(2 * n) * QW * 4));
  PrepArgs pa{};
  pa.codes = d.codes.as<uint8_t>();
  pa.code_stride = (int)code_stride;
  pa.qual = has_qual ? d.qual.as<uint8_t>() : nullptr;
  pa.qual_stride = (int)qual_stride;
  pa.first_qual = has_first_qual ? d.first_qual.as<uint8_t>() : nullptr;
  pa.len = d.len.as<int32_t>();
  pa.n = n;
  pa.color = m->p.color;
  pa.planes = d.planes.as<uint32_t>();
  pa.qplanes = d.qplanes.as<uint32_t>();
  HIPOK(launch_prep_reads(pa, m->st));
  return STORM_OK;
}

static float elapsed(storm_mapper* m, hipEvent_t a, hipEvent_t b) {
  float ms = 0;
  (void)hipEventElapsedTime(&ms, a, b);
  return ms;
}

// PASS_HIT_LIMIT, or the value of STORM_PASS_HIT_LIMIT in the environment (read once; tests use a
// small value to exercise the splitting)
static int64_t pass_hit_limit() {
  static const int64_t v = [] {
    const char* e = getenv("STORM_PASS_HIT_LIMIT");
    const long long x = e ? atoll(e) : 0;
    return x > 0 ? (int64_t)x : PASS_HIT_LIMIT;
  }();
  return v;
}

// Reads [r0, r1) in passes of at most `chunk` reads, each halved while its raw hits exceed
// pass_hit_limit().
int run_passes(storm_mapper* m, DevReads& dr, Slot* slots, const CellTarget& ct, int ref_id, int& pass_id, int64_t r0,
               int64_t r1, PassStats& stats, int64_t chunk) {
  for (int64_t c0 = r0; c0 < r1;) {
    const int64_t c1 = std::min<int64_t>(r1, c0 + chunk);
    const int rc = run_pass(m, dr, slots, ct, ref_id, pass_id++, c0, c1, stats);
    if (rc == PASS_SPLIT) {
      chunk = std::max<int64_t>(1, (c1 - c0) / 2);
      continue;
    }
    if (rc) return rc;
    c0 = c1;
  }
  return STORM_OK;
}

// One mapping pass of reads [r0, r1) (device order) against one reference sequence: the stages
// of process_read (run.c:555-747) as separate launches, with host synchronisations to size
// the buffers of the next stage.
int run_pass(storm_mapper* m, DevReads& dr, Slot* slots, const CellTarget& ct, int ref_id, int pass_id, int64_t r0,
             int64_t r1, PassStats& stats) {
  PassScratch& P = m->ps;
  hipStream_t st = m->st;
  RefState& rs = *m->res->rstate[(size_t)ref_id];
  const int64_t n_rs = 2 * (r1 - r0);
  if (n_rs == 0) return STORM_OK;
  ScoreDev sc{};
  fill_score(m->p, sc);
  hipEvent_t e[7];
  for (auto& x : e) HIPOK(hipEventCreate(&x));
  struct EvGuard { hipEvent_t* e; ~EvGuard() { for (int i = 0; i < 7; ++i) (void)hipEventDestroy(e[i]); } } eg{e};

  SeedArgs sa{};
  sa.r = dr.view();
  sa.seeds = seed_set(rs);
  sa.I = m->p.allowed_indels;
  sa.sense_mask = m->p.sense_mask;
  sa.rs_begin = 2 * r0;
  sa.n_rs = n_rs;
  HIPOK(P.seg_count.ensure((size_t)(n_rs + 1) * 8));
  HIPOK(P.seg_off.ensure((size_t)(n_rs + 1) * 8));
  HIPOK(P.cand_off.ensure((size_t)(n_rs + 1) * 8));
  HIPOK(P.overflow.ensure((size_t)n_rs));
  HIPOK(P.n_overflow.ensure(4));
  HIPOK(hipMemsetAsync(P.seg_count.p, 0, (size_t)(n_rs + 1) * 8, st));
  HIPOK(hipMemsetAsync(P.n_overflow.p, 0, 4, st));
  sa.seg_count = P.seg_count.as<int64_t>();
  HIPOK(hipEventRecord(e[0], st));
  HIPOK(launch_seed_count(sa, st));   // the index lookups alone: first[key], first[key + 1]
  HIPOK(hipEventRecord(e[6], st));
  size_t tmpb = 0;
  HIPOK(scan_exclusive_i64(nullptr, nullptr, n_rs + 1, nullptr, tmpb, st));
  size_t tmpb2 = 0;
  HIPOK(scan_exclusive_u8(nullptr, nullptr, 1 << 20, nullptr, tmpb2, st));
  HIPOK(P.cub_tmp.ensure(std::max(tmpb, tmpb2) + (1 << 20)));
  tmpb = P.cub_tmp.n;
  HIPOK(scan_exclusive_i64(P.seg_count.as<int64_t>(), P.seg_off.as<int64_t>(), n_rs + 1, P.cub_tmp.p, tmpb, st));
  int64_t total = 0;
  HIPOK(hipMemcpyAsync(&total, P.seg_off.as<int64_t>() + n_rs, 8, hipMemcpyDeviceToHost, st));
  HIPOK(hipStreamSynchronize(st));
  // the device scans and sorts count items in 32 bits and the per-pass buffers grow with the raw
  // hits: a pass over more than PASS_HIT_LIMIT hits is split by the caller (nothing written yet)
  const int64_t limit = pass_hit_limit();
  if (total > limit) {
    if (r1 - r0 > 1) return PASS_SPLIT;
    return fail(STORM_ERR_MEMORY, "read %lld has %lld raw index hits (more than %lld); filter repeats with a lower -l",
                (long long)r0, (long long)total, (long long)limit);
  }
  stats.hits += total;
  HIPOK(P.hits.ensure((size_t)std::max<int64_t>(total, 1) * 4));
  HIPOK(P.kept.ensure((size_t)std::max<int64_t>(total, 1) * 4));
  sa.seg_off = P.seg_off.as<int64_t>();
  sa.hits = P.hits.as<int32_t>();
  sa.kept = P.kept.as<int32_t>();
  sa.overflow = P.overflow.as<uint8_t>();
  sa.n_overflow = P.n_overflow.as<int32_t>();
  HIPOK(launch_seed_gather(sa, st));
  int32_t n_over = 0;
  HIPOK(hipMemcpyAsync(&n_over, P.n_overflow.p, 4, hipMemcpyDeviceToHost, st));
  HIPOK(hipStreamSynchronize(st));
  if (n_over > 0) {   // segments beyond LDS: device segmented radix sort, then the keep rule
    HIPOK(P.seg_begin.ensure((size_t)n_rs * 8));
    HIPOK(P.seg_end.ensure((size_t)n_rs * 8));
    HIPOK(P.sorted.ensure((size_t)std::max<int64_t>(total, 1) * 4));
    HIPOK(launch_overflow_ranges(P.seg_off.as<int64_t>(), P.overflow.as<uint8_t>(), n_rs, P.seg_begin.as<int64_t>(),
                                 P.seg_end.as<int64_t>(), st));
    size_t sb = 0;
    HIPOK(segsort_i32(nullptr, nullptr, total, n_rs, nullptr, nullptr, nullptr, sb, st));
    HIPOK(P.cub_tmp.ensure(sb + 256));
    sb = P.cub_tmp.n;
    HIPOK(segsort_i32(P.hits.as<int32_t>(), P.sorted.as<int32_t>(), total, n_rs, P.seg_begin.as<int64_t>(),
                      P.seg_end.as<int64_t>(), P.cub_tmp.p, sb, st));
    SeedArgs su = sa;
    su.hits = P.sorted.as<int32_t>();
    HIPOK(launch_unique_segments(su, st));
    HIPOK(hipStreamSynchronize(st));
  }
  // candidate offsets and the dense list
  HIPOK(hipMemsetAsync(P.seg_count.as<int64_t>() + n_rs, 0, 8, st));
  tmpb = P.cub_tmp.n;
  HIPOK(scan_exclusive_i64(P.seg_count.as<int64_t>(), P.cand_off.as<int64_t>(), n_rs + 1, P.cub_tmp.p, tmpb, st));
  int64_t n_cand = 0;
  HIPOK(hipMemcpyAsync(&n_cand, P.cand_off.as<int64_t>() + n_rs, 8, hipMemcpyDeviceToHost, st));
  HIPOK(hipStreamSynchronize(st));
  stats.cand += n_cand;
  const size_t cb = (size_t)std::max<int64_t>(n_cand, 1);
  HIPOK(P.cand_rs.ensure(cb * 4));
  HIPOK(P.cand_diag.ensure(cb * 4));
  HIPOK(P.cand_pos.ensure(cb * 4));
  HIPOK(P.pass.ensure(cb + 16));
  HIPOK(P.pass_scan.ensure((cb + 1) * 8));
  CompactArgs ca{};
  ca.seg_off = P.seg_off.as<int64_t>();
  ca.seg_count = P.seg_count.as<int64_t>();
  ca.cand_off = P.cand_off.as<int64_t>();
  ca.kept = P.kept.as<int32_t>();
  ca.rs_begin = 2 * r0;
  ca.n_rs = n_rs;
  ca.d = sc.d;
  ca.cand_rs = P.cand_rs.as<int32_t>();
  ca.cand_diag = P.cand_diag.as<int32_t>();
  ca.cand_pos = P.cand_pos.as<int32_t>();
  HIPOK(launch_compact_candidates(ca, st));
  HIPOK(hipEventRecord(e[1], st));
  // filter
  FilterArgs fa{};
  fa.g = rs.g.view();
  fa.rplanes = dr.planes.as<uint32_t>();
  fa.rlen = dr.len.as<int32_t>();
  fa.len_shift = 1;
  fa.cand_read = P.cand_rs.as<int32_t>();
  fa.cand_pos = P.cand_pos.as<int32_t>();
  fa.n = n_cand;
  fa.score = nullptr;
  fa.pass = P.pass.as<uint8_t>();
  fa.mt = sc.fmatch; fa.mm = sc.fmismatch; fa.op = sc.fopen; fa.ex = sc.fext; fa.thr = sc.fthr;
  HIPOK(launch_filter(fa, sc.W, st));
  HIPOK(hipEventRecord(e[2], st));
  // records
  HIPOK(hipMemsetAsync(P.pass.as<uint8_t>() + n_cand, 0, 1, st));
  size_t tb3 = 0;
  HIPOK(scan_exclusive_u8(nullptr, nullptr, n_cand + 1, nullptr, tb3, st));
  HIPOK(P.cub_tmp.ensure(tb3 + 256));
  tb3 = P.cub_tmp.n;
  HIPOK(scan_exclusive_u8(P.pass.as<uint8_t>(), P.pass_scan.as<int64_t>(), n_cand + 1, P.cub_tmp.p, tb3, st));
  int64_t n_rec = 0;
  HIPOK(hipMemcpyAsync(&n_rec, P.pass_scan.as<int64_t>() + n_cand, 8, hipMemcpyDeviceToHost, st));
  HIPOK(hipStreamSynchronize(st));
  stats.rec += n_rec;
  const size_t rb = (size_t)std::max<int64_t>(n_rec, 1);
  HIPOK(P.rec_rs.ensure(rb * 4));
  HIPOK(P.rec_diag.ensure(rb * 4));
  HIPOK(P.rec_off.ensure((size_t)(n_rs + 1) * 8));
  HIPOK(P.recs.ensure(rb * sizeof(AlignRec)));
  HIPOK(P.tb.ensure(rb * TBB));
  RecArgs ra{};
  ra.pass = P.pass.as<uint8_t>();
  ra.pass_scan = P.pass_scan.as<int64_t>();
  ra.cand_rs = P.cand_rs.as<int32_t>();
  ra.cand_diag = P.cand_diag.as<int32_t>();
  ra.n_cand = n_cand;
  ra.cand_off = P.cand_off.as<int64_t>();
  ra.n_rs = n_rs;
  ra.rec_rs = P.rec_rs.as<int32_t>();
  ra.rec_diag = P.rec_diag.as<int32_t>();
  ra.rec_off = P.rec_off.as<int64_t>();
  HIPOK(launch_make_records(ra, st));
  // DP + traceback, chunked by the scratch budget
  AlignArgs aa{};
  aa.g = rs.g.view();
  aa.r = dr.view();
  aa.sc = sc;
  aa.rec_rs = P.rec_rs.as<int32_t>();
  aa.rec_diag = P.rec_diag.as<int32_t>();
  aa.out = P.recs.as<AlignRec>();
  aa.tb = P.tb.as<uint8_t>();
  int maxL = 0;
  {
    std::vector<int32_t> lens((size_t)(r1 - r0));
    HIPOK(hipMemcpyAsync(lens.data(), dr.len.as<int32_t>() + r0, lens.size() * 4, hipMemcpyDeviceToHost, st));
    HIPOK(hipStreamSynchronize(st));
    for (int32_t x : lens) maxL = std::max(maxL, (int)x);
  }
  {
    const int rc = align_records(aa, maxL, n_rec, P.align_scratch, st);
    if (rc) return rc;
  }
  HIPOK(hipEventRecord(e[3], st));
  // hit map
  ReplayArgs pa{};
  pa.rec = P.recs.as<AlignRec>();
  pa.rec_off = P.rec_off.as<int64_t>();
  pa.len = dr.len.as<int32_t>();
  pa.read_begin = r0;
  pa.n_reads = r1 - r0;
  pa.ref_id = ref_id;
  pa.pass = pass_id;
  pa.t = m->p.t;
  pa.sense_mask = m->p.sense_mask;
  pa.slots = slots;
  HIPOK(launch_replay(pa, st));
  HIPOK(hipEventRecord(e[4], st));
  // the tracebacks this pass inserted go to the reads' cell store (device to device)
  if (m->want_tb)
    HIPOK(launch_store_cells(slots, 4 * r0, 4 * (r1 - r0), pass_id, tb_ranks(m->p), P.tb.as<uint8_t>(), ct.cells, ct.ids,
                             st));
  HIPOK(hipEventRecord(e[5], st));
  HIPOK(hipEventSynchronize(e[5]));
  stats.ms_seed += elapsed(m, e[0], e[1]);
  stats.ms_lookup += elapsed(m, e[0], e[6]);
  stats.ms_filter += elapsed(m, e[1], e[2]);
  stats.ms_align += elapsed(m, e[2], e[3]);
  stats.ms_replay += elapsed(m, e[3], e[4]);
  return STORM_OK;
}

// whole map call: reads of the view in the given order; results per read in view order

// seed-key lookups of a read set: per strand in the sense mask, per seed, its read positions
static int64_t count_lookups(const storm_mapper* m, const storm_reads_view* v);

static_assert(sizeof(storm_hit) == sizeof(HitRec) && offsetof(storm_hit, tb_len) == offsetof(HitRec, tb_len),
              "storm_hit mirrors HitRec");

// The end of a map call: the device slots (device order) and the reads' traceback cells go to the
// caller's arrays (ho: four int32 arrays; ro: whole records and tracebacks) or to host slots in
// view order (SAM output).  Returns the mapped reads.
// Bytes of a traceback record on its way to the caller: a path through a band of 2I + 1 columns
// takes at most L + I steps (one per read code, plus at most I horizontal gap steps, the indel
// cap), 2 codes per byte, so records of the call's longest read fit in (maxL + I + 1) / 2 bytes,
// rounded up to 16; the caller's records are TBB bytes, the rest zero.  The environment variable
// STORM_TB_FULL_WIDTH copies whole TBB-byte records instead (for tests).
static int tb_width(const storm_mapper* m, int maxL = -1) {
  if (getenv("STORM_TB_FULL_WIDTH")) return TBB;
  if (maxL < 0) {   // the call's longest read, from its length histogram
    maxL = 0;
    for (int L = MAXL + 1; L > 0 && !maxL; --L)
      if ((size_t)L < m->len_hist.size() && m->len_hist[(size_t)L]) maxL = L;
  }
  if (maxL <= 0 || maxL > MAXL) return TBB;
  const int I = std::max(0, std::min(m->p.allowed_indels, MAX_I));
  const int w = ((maxL + I + 1) / 2 + 15) / 16 * 16;
  return std::min(w, TBB);
}

int finish_view(storm_mapper* m, const Slot* slots, const uint8_t* cells, const std::vector<int64_t>& order, int64_t n,
                const HitsOut* ho, const RecordsOut* ro, std::vector<HostSlot>& out_slots, int64_t& mapped) {
  if (ho) return hits_to_host(m, slots, order, n, *ho, mapped);
  const int M = tb_ranks(m->p);
  const bool want_tb = m->want_tb && (!ro || ro->tb);
  bool identity = true;
  for (int64_t i = 0; i < n && identity; ++i) identity = order[(size_t)i] == i;
  if (!identity && m->order_dev_n != n) {   // the upload left no order on the device
    std::vector<int32_t> o32((size_t)n);
    for (int64_t i = 0; i < n; ++i) o32[(size_t)i] = (int32_t)order[(size_t)i];
    HIPOK(m->order_dev.ensure((size_t)n * 4));
    HIPOK(hipMemcpyAsync(m->order_dev.p, o32.data(), (size_t)n * 4, hipMemcpyHostToDevice, m->st));
    HIPOK(hipStreamSynchronize(m->st));
    m->order_dev_n = n;
  }
  const int tw = ro ? tb_width(m) : TBB;   // records for the caller leave compact
  HIPOK(m->rec_dev.ensure((size_t)(4 * n) * sizeof(HitRec)));
  if (want_tb) HIPOK(m->tb_dev.ensure((size_t)(n * M) * tw));
  HIPOK(m->mapped_dev.ensure(16));
  HIPOK(hipMemsetAsync(m->mapped_dev.p, 0, 4, m->st));
  HIPOK(launch_hit_records(slots, identity ? nullptr : m->order_dev.as<int32_t>(), 0, n, M, cells, m->rec_dev.as<HitRec>(),
                           want_tb ? m->tb_dev.as<uint8_t>() : nullptr, tw, m->mapped_dev.as<int>(), m->st));
  int mapped_d = 0;
  HIPOK(hipMemcpyAsync(&mapped_d, m->mapped_dev.p, 4, hipMemcpyDeviceToHost, m->st));
  std::vector<storm_hit> hv;
  std::vector<uint8_t> tv;
  storm_hit* hits = ro ? ro->hits : nullptr;
  uint8_t* tb = ro ? ro->tb : nullptr;
  if (!ro) {
    hv.resize((size_t)(4 * n));
    hits = hv.data();
    if (want_tb) { tv.resize((size_t)(n * M) * TBB); tb = tv.data(); }
  }
  std::vector<CopySeg> out = {{hits, m->rec_dev.p, (size_t)(4 * n) * sizeof(HitRec)}};
  if (want_tb) out.push_back({tb, m->tb_dev.p, (size_t)(n * M) * tw, (size_t)tw, (size_t)TBB});
  const int rc = staged_copy(m, out, true);   // the count came with the stream's first copy
  if (rc) return rc;
  mapped = mapped_d;
  if (ro) return STORM_OK;
  out_slots.assign((size_t)(4 * n), HostSlot{});
  for (int64_t q = 0; q < 4 * n; ++q) {
    const storm_hit& h = hits[q];
    HostSlot& o = out_slots[(size_t)q];
    o.score = h.score; o.ref_id = h.ref_id; o.ref_start = h.ref_start; o.ref_end = h.ref_end;
    o.read_start = h.read_start; o.read_end = h.read_end; o.sense = h.sense; o.tb_len = h.tb_len;
    const int k = (int)(q & 3);
    if (want_tb && h.ref_id != -1 && k < M) {
      const uint8_t* src = tb + ((q >> 2) * M + k) * (int64_t)TBB;
      o.tb.assign(src, src + TBB);
    }
  }
  return STORM_OK;
}

static int map_view_staged(storm_mapper* m, const storm_reads_view* v, int mode, std::vector<HostSlot>& out_slots,
                           storm_map_stats* st_out, const HitsOut* ho, const RecordsOut* ro, const FinishFn* fin) {
  const int64_t n = v->n_reads;
  if (n <= 0) return fail(STORM_ERR_INVALID, "no reads");
  // plain mode: equal-length batches (the reference maps one read length per run); the device
  // order groups reads by length, stable within a length
  std::vector<int64_t>& order = m->order64;
  bool sorted_view = true;
  length_order_chunks(v, {0, n}, order, sorted_view, m->len_hist, &m->order32);
  std::vector<std::pair<int64_t, int64_t>> batches;
  if (mode == STORM_MODE_PLAIN) {
    int64_t s = 0;
    for (int64_t i = 1; i <= n; ++i)
      if (i == n || v->len[order[(size_t)i]] != v->len[order[(size_t)s]]) { batches.emplace_back(s, i); s = i; }
  } else {
    batches.emplace_back(0, n);
  }
  const double t0 = now_s();
  DevReads& dr = m->plain_reads;   // kept across calls: a call allocates only when it outgrows them
  int rc = upload_reads(m, v, order, dr, &m->len_hist, &sorted_view, &m->order32);
  if (rc) return rc;
  DevMem& slots = m->plain_slots;
  HIPOK(slots.ensure((size_t)(4 * n) * sizeof(Slot)));
  if (m->want_tb) HIPOK(m->plain_cells.ensure((size_t)(4 * n) * TBB));
  HIPOK(launch_slots_init(slots.as<Slot>(), 4 * n, m->st));
  const CellTarget ct{m->plain_cells.as<uint8_t>(), nullptr};
  PassStats ps;
  int pass_id = 0;
  const double ti = m->res->index_seconds;
  for (size_t r = 0; r < m->res->refs.size(); ++r) {
    rc = ensure_ref_resident(m, r);
    if (rc) return rc;
    for (auto& b : batches) {
      // bound the per-pass buffers: at most 1M reads per launch group
      rc = run_passes(m, dr, slots.as<Slot>(), ct, (int)r, pass_id, b.first, b.second, ps, 1 << 20);
      if (rc) return rc;
    }
    if (m->res->refs.size() > 1 && !m->res->pinned) release_ref(m, r);
  }
  int64_t mapped = 0;
  const double tf = now_s();
  if (fin) rc = (*fin)(FinishCtx{slots.as<Slot>(), m->plain_cells.as<uint8_t>(), &dr, &order, n}, mapped);
  else rc = finish_view(m, slots.as<Slot>(), m->plain_cells.as<uint8_t>(), order, n, ho, ro, out_slots, mapped);
  if (rc) return rc;
  m->finish_seconds = fin ? now_s() - tf : 0;
  if (st_out) {
    st_out->n_reads = n;
    st_out->n_mapped = mapped;
    st_out->n_candidates = ps.cand;
    st_out->n_filter_pass = ps.rec;
    st_out->n_accepted = 0;
    st_out->n_hits = ps.hits;
    st_out->seconds_index = m->res->index_seconds - ti;
    st_out->seconds_map = now_s() - t0 - st_out->seconds_index - m->finish_seconds;
    st_out->gpu_ms_seed = ps.ms_seed;
    st_out->gpu_ms_filter = ps.ms_filter;
    st_out->gpu_ms_align = ps.ms_align;
    st_out->gpu_ms_replay = ps.ms_replay;
    st_out->gpu_ms_lookup = ps.ms_lookup;
  }
  return STORM_OK;
}

// ------------------------------------------------------------------------------ batched path
static int device_cus(int device) {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return 256;
  return prop.multiProcessorCount;
}

static int ensure_fused_buffers(storm_mapper* m, int64_t n, int64_t stride, int grid) {
  FusedState& F = m->fs;
  if (n > F.cap_reads) {
    const int64_t cap = std::max<int64_t>(n, F.cap_reads);
    HIPOK(F.slots.ensure((size_t)(4 * cap) * sizeof(Slot)));
    HIPOK(F.cells.ensure((size_t)(4 * cap) * TBB));
    HIPOK(F.overflow.ensure((size_t)cap));
    HIPOK(F.gids.ensure((size_t)(4 * cap) * 4));
    HIPOK(F.gtb.ensure((size_t)(4 * cap) * TBB));
    F.cap_reads = cap;
  }
  if (stride > F.scratch_stride || grid > F.grid) {
    const int64_t st = std::max(stride, F.scratch_stride);
    const int g = std::max(grid, F.grid);
    HIPOK(F.scratch.ensure((size_t)g * 64 * st));
    HIPOK(F.tbbuf.ensure((size_t)g * FUSED_PCAP * TBB));
    HIPOK(F.work.ensure((size_t)g * FUSED_WORK_BYTES));
    F.scratch_stride = st;
    F.grid = g;
  }
  HIPOK(F.n_dev.ensure(16));
  HIPOK(F.counter.ensure(16));
  HIPOK(F.stats.ensure(64));
  HIPOK(F.gcount.ensure(16));
  return STORM_OK;
}

static FusedArgs fused_args(storm_mapper* m, size_t ref, int group) {
  FusedState& F = m->fs;
  FusedArgs a{};
  a.g = m->res->rstate[ref]->g.view();
  a.r = F.reads.view();
  a.seeds = seed_set(*m->res->rstate[ref]);
  fill_score(m->p, a.sc);
  a.n_reads = F.n_dev.as<int32_t>();
  a.group = group;
  a.task_counter = F.counter.as<int32_t>();
  a.slots = F.slots.as<Slot>();
  a.cells = F.cells.as<uint8_t>();
  a.overflow = F.overflow.as<uint8_t>();
  a.scratch = F.scratch.as<uint8_t>();
  a.scratch_stride = F.scratch_stride;
  a.tbbuf = F.tbbuf.as<uint8_t>();
  a.work = F.work.as<uint8_t>();
  a.ref_id = (int)ref;
  a.sense_mask = m->p.sense_mask;
  a.stats = F.stats.as<unsigned long long>();
  a.phase = F.phase.as<unsigned long long>();
  return a;
}

// Autotuner: reads per task and persistent grid for this band class and length bucket, timed on
// a sample of the batch (the sample's slots are reset afterwards).
// the tuner's key: band class and 32-nt length bucket of the longest read
static int64_t tune_key(const storm_mapper* m, int maxL) {
  ScoreDev sc{};
  fill_score(m->p, sc);
  return ((int64_t)sc.W << 40) | ((int64_t)sc.I << 32) | ((int64_t)m->p.color << 24) | (int64_t)((maxL + 31) / 32);
}

// a tuning of this context or, on the same device, of a sibling context
static bool cached_tune(storm_mapper* m, int64_t key, FusedTune& out) {
  FusedState& F = m->fs;
  auto it = F.tune.find(key);
  if (it != F.tune.end()) { out = it->second; return true; }
  if (m->shared_tune) {
    std::lock_guard<std::mutex> lk(m->shared_tune->mu);
    auto s = m->shared_tune->best.find(key);
    if (s != m->shared_tune->best.end()) {
      F.tune[key] = s->second;
      out = s->second;
      return true;
    }
  }
  return false;
}

static int tune_fused(storm_mapper* m, int64_t n, int maxL, int max_grid, FusedTune& out) {
  FusedState& F = m->fs;
  const int64_t key = tune_key(m, maxL);
  if (cached_tune(m, key, out)) return STORM_OK;
  const int64_t sample = std::min<int64_t>(n, 32768);
  const int32_t ns[2] = {(int32_t)sample, 0};
  HIPOK(hipMemcpyAsync(F.n_dev.p, ns, 8, hipMemcpyHostToDevice, m->st));
  // the two timing events, destroyed on every return
  struct Events {
    hipEvent_t e[2] = {nullptr, nullptr};
    ~Events() {
      for (hipEvent_t x : e)
        if (x) (void)hipEventDestroy(x);
    }
  } ev;
  HIPOK(hipEventCreate(&ev.e[0]));
  HIPOK(hipEventCreate(&ev.e[1]));
  hipEvent_t e0 = ev.e[0], e1 = ev.e[1];
  FusedTune best;
  best.ms = 1e30f;
  {   // warm-up launch (code object load, caches)
    FusedArgs a = fused_args(m, 0, 16);
    HIPOK(launch_fused_reset(a.task_counter, a.stats, m->st));
    HIPOK(launch_fused(a, max_grid, m->st));
    HIPOK(hipStreamSynchronize(m->st));
  }
  auto timed = [&](int G, int grid, float& ms) -> int {
    FusedArgs a = fused_args(m, 0, G);
    HIPOK(launch_slots_init(F.slots.as<Slot>(), 4 * sample, m->st));
    HIPOK(launch_fused_reset(a.task_counter, a.stats, m->st));
    HIPOK(hipEventRecord(e0, m->st));
    HIPOK(launch_fused(a, grid, m->st));
    HIPOK(hipEventRecord(e1, m->st));
    HIPOK(hipEventSynchronize(e1));
    HIPOK(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best.ms) { best.ms = ms; best.group = G; best.grid = grid; }
    return STORM_OK;
  };
  // reads per task G = 8, 16, 32, 64 on the occupancy-sized grid, stopping at the first G slower
  // than the one before (fewer, longer tasks leave workgroups idle on the sample: a larger G was
  // never faster once one was slower, and such runs take several times the best one), then the
  // best G on half the grid
  float prev = 1e30f;
  for (int G = 8; G <= 64; G *= 2) {
    float ms = 0;
    int rc = timed(G, max_grid, ms);
    if (rc) return rc;
    if (ms > prev) break;
    prev = ms;
  }
  if (max_grid / 2 >= 1) {
    float ms = 0;
    int rc = timed(best.group, max_grid / 2, ms);
    if (rc) return rc;
  }
  HIPOK(launch_slots_init(F.slots.as<Slot>(), 4 * sample, m->st));
  F.tune[key] = best;
  if (m->shared_tune) {
    std::lock_guard<std::mutex> lk(m->shared_tune->mu);
    m->shared_tune->best.emplace(key, best);
  }
  out = best;
  return STORM_OK;
}

// Reads flagged by the fused kernel (LDS capacity exceeded) go through the staged kernels for
// this reference: their planes and slots are gathered on the device into a sub-batch, mapped,
// and the slots scattered back.
int staged_subset(storm_mapper* m, const std::vector<int64_t>& dev_ids, int ref, int& pass_id, PassStats& ps) {
  FusedState& F = m->fs;
  const int64_t k = (int64_t)dev_ids.size();
  std::vector<int32_t> ids((size_t)k);
  for (int64_t i = 0; i < k; ++i) ids[(size_t)i] = (int32_t)dev_ids[(size_t)i];
  DevReads& dr = m->sub_reads;   // grown when a fallback outgrows it, kept across calls
  HIPOK(m->sub_ids.ensure((size_t)k * 4));
  HIPOK(hipMemcpyAsync(m->sub_ids.p, ids.data(), (size_t)k * 4, hipMemcpyHostToDevice, m->st));
  HIPOK(dr.planes.ensure((size_t)(2 * k) * 2 * RW * 4));
  HIPOK(dr.qplanes.ensure((size_t)(2 * k) * QW * 4));
  HIPOK(dr.len.ensure((size_t)k * 4));
  HIPOK(m->sub_slots.ensure((size_t)(4 * k) * sizeof(Slot)));
  dr.n = k;
  dr.has_qual = F.reads.has_qual;
  const int32_t* d_ids = m->sub_ids.as<int32_t>();
  Slot* sl = m->sub_slots.as<Slot>();
  HIPOK(launch_subset_gather(F.reads.planes.as<uint32_t>(), F.reads.qplanes.as<uint32_t>(), F.reads.len.as<int32_t>(),
                             F.slots.as<Slot>(), d_ids, k, dr.planes.as<uint32_t>(), dr.qplanes.as<uint32_t>(),
                             dr.len.as<int32_t>(), sl, m->st));
  const CellTarget ct{F.cells.as<uint8_t>(), d_ids};   // the subset's cells are the batch's
  int rc = run_passes(m, dr, sl, ct, ref, pass_id, 0, k, ps, k);
  if (rc) return rc;
  HIPOK(launch_subset_scatter(sl, d_ids, k, F.slots.as<Slot>(), m->st));
  HIPOK(hipStreamSynchronize(m->st));   // ids is released on return
  return STORM_OK;
}

// One graph per reference: reset + fused kernel, replayed for every batch with the same buffers
// and scoring (the kernel arguments are part of the captured graph; the read range is read from
// the device).  sig is the graph's key in the mapper's cache.
static int fused_graph(storm_mapper* m, size_t r, const FusedTune& tune, std::vector<const void*>& sig,
                       hipGraphExec_t& exec) {
  FusedState& F = m->fs;
  FusedArgs a = fused_args(m, r, tune.group);
  sig = {a.g.lh, a.r.planes, a.r.qual, a.r.len, a.slots, a.cells, a.scratch, a.tbbuf, a.work,
         a.seeds.s[0].first, (const void*)(intptr_t)tune.group, (const void*)(intptr_t)tune.grid,
         (const void*)(intptr_t)a.scratch_stride, (const void*)(intptr_t)a.sense_mask};
  const int* scw = (const int*)&a.sc;
  for (size_t q = 0; q < sizeof(ScoreDev) / sizeof(int); ++q) sig.push_back((const void*)(intptr_t)scw[q]);
  sig.push_back((const void*)r);
  exec = F.find_graph(sig);
  if (!exec) {
    hipGraph_t g;
    HIPOK(hipStreamBeginCapture(m->st, hipStreamCaptureModeThreadLocal));
    hipError_t e1 = launch_fused_reset(a.task_counter, a.stats, m->st);
    hipError_t e2 = launch_fused(a, tune.grid, m->st);
    HIPOK(hipStreamEndCapture(m->st, &g));
    HIPOK(e1);
    HIPOK(e2);
    HIPOK(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
    (void)hipGraphDestroy(g);
    F.add_graph(sig, exec);
  }
  return STORM_OK;
}

int fused_core(storm_mapper* m, int64_t n, int maxL, const FallbackFn& fallback, PassStats& ps,
               unsigned long long totals[4], int64_t& n_fallback) {
  FusedState& F = m->fs;
  ScoreDev sc{};
  fill_score(m->p, sc);
  const int64_t stride = (int64_t)maxL * dp_dir_words(sc.I) * 4;   // direction words per lane
  const int bpc = fused_blocks_per_cu(sc.W, sc.I, m->p.color);
  if (bpc <= 0) return fail(STORM_ERR_DEVICE, "fused kernel: no occupancy for this band class");
  const int max_grid = device_cus(m->device) * bpc;
  int rc = ensure_fused_buffers(m, n, stride, max_grid);
  if (rc) return rc;
  HIPOK(launch_slots_init(F.slots.as<Slot>(), 4 * n, m->st));
#ifdef STORM_FUSED_PHASES
  if (!F.phase.p) {
    HIPOK(F.phase.ensure(FUSED_PHASES * sizeof(unsigned long long)));
    HIPOK(hipMemsetAsync(F.phase.p, 0, FUSED_PHASES * sizeof(unsigned long long), m->st));
  }
#endif
  const int32_t n32[2] = {(int32_t)n, 0};   // reads [0, n)
  const size_t n_refs = m->res->refs.size();
  int pass_id = 0;
  for (int q = 0; q < 4; ++q) totals[q] = 0;
  n_fallback = 0;
  std::vector<uint8_t>& ovf = m->ovf_host;
  if ((int64_t)ovf.size() < n) ovf.resize((size_t)n);
  for (size_t r = 0; r < n_refs; ++r) {
    rc = ensure_ref_resident(m, r);
    if (rc) return rc;
    FusedTune tune;
    if (r == 0) {
      rc = tune_fused(m, n, maxL, max_grid, tune);
      if (rc) return rc;
    } else {
      tune_fused(m, n, maxL, max_grid, tune);
    }
    HIPOK(hipMemcpyAsync(F.n_dev.p, n32, 8, hipMemcpyHostToDevice, m->st));
    std::vector<const void*> sig;
    hipGraphExec_t exec = nullptr;
    rc = fused_graph(m, r, tune, sig, exec);
    if (rc) return rc;
#ifdef STORM_FUSED_PHASES
    unsigned long long ph0[FUSED_PHASES], ph1[FUSED_PHASES];
    HIPOK(hipMemcpyAsync(ph0, F.phase.p, sizeof ph0, hipMemcpyDeviceToHost, m->st));
    HIPOK(hipStreamSynchronize(m->st));
#endif
    if (!m->ev0) HIPOK(hipEventCreate(&m->ev0));
    if (!m->ev1) HIPOK(hipEventCreate(&m->ev1));
    hipEvent_t e0 = m->ev0, e1 = m->ev1;
    HIPOK(hipEventRecord(e0, m->st));
    HIPOK(hipGraphLaunch(exec, m->st));
    HIPOK(hipEventRecord(e1, m->st));
    unsigned long long st3[4];
    HIPOK(hipMemcpyAsync(st3, F.stats.p, sizeof st3, hipMemcpyDeviceToHost, m->st));
    HIPOK(hipMemcpyAsync(ovf.data(), F.overflow.p, (size_t)n, hipMemcpyDeviceToHost, m->st));
    HIPOK(hipStreamSynchronize(m->st));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    ps.ms_filter += ms;   // the fused kernel carries every stage
#ifdef STORM_FUSED_PHASES
    {   // wavefront-time shares of the launch's phases (constant 100 MHz clock)
      HIPOK(hipMemcpy(ph1, F.phase.p, sizeof ph1, hipMemcpyDeviceToHost));
      double tot = 0, d[FUSED_PHASES];
      for (int q = 0; q < FUSED_PHASES; ++q) { d[q] = (double)(ph1[q] - ph0[q]); tot += d[q]; }
      fprintf(stderr, "fused phases W=%d I=%d color=%d reads=%lld grid=%d G=%d launch=%.2f ms wave-s=%.3f: candidates %.1f%%, "
              "filter %.1f%%, scan+triage %.1f%%, dp+traceback %.1f%%, replay+cells %.1f%%, task %.1f%%\n",
              sc.W, sc.I, m->p.color, (long long)n, tune.grid, tune.group, ms, tot * 1e-8, 100 * d[0] / tot,
              100 * d[1] / tot, 100 * d[2] / tot, 100 * d[3] / tot, 100 * d[4] / tot, 100 * d[5] / tot);
    }
#endif
    for (int q = 0; q < 4; ++q) totals[q] += st3[q];
    std::vector<int64_t> over;
    for (int64_t i = 0; i < n; ++i) if (ovf[(size_t)i]) over.push_back(i);
    n_fallback += (int64_t)over.size();
    if (!over.empty()) {
      rc = fallback(over, (int)r, pass_id, ps);
      if (rc) return rc;
    }
    if (n_refs > 1 && !m->res->pinned) {
      release_ref(m, r);
      F.drop_graph(sig);   // its genome and index buffers are gone
    }
  }
  return STORM_OK;
}

static int batched_stats(storm_mapper* m, int64_t n, int64_t mapped, const unsigned long long totals[4],
                         const PassStats& ps, int64_t n_fallback, double t0, double ti, storm_map_stats* st_out);

// Chunks of a batched records call.  Splitting a launch in K hides (K - 1) / K of the results'
// copy (measured 36-60 GB/s through the pinned halves) and adds K - 1 launch tails (the last tasks
// of a persistent launch run with the chip partly idle: measured 0.4-1.0 times the tuner's sample
// launch, one task per workgroup), so K = sqrt(copy / tail) with the tail taken at its upper
// figure, rounded down; the tuner's figure is known from the first call of a band class on, which
// itself runs in one launch.  Measured with whole 144-byte traceback records: 5M color reads of
// 34-74 on 50 Mbp gained in 2 chunks (host share 73 -> 51 ms, launches 411 -> 418 ms), while 4M
// x 100 nt at -i 7 -M 4 (a 25 ms tail against 60 ms of copies) and 100k x 50 nt (two launches
// 0.26 ms longer than one, copies 0.45 ms) did not; with the compact records (tb_width) the copies
// are small enough that all three run in one launch.  The measurement build with phase timers
// always maps in one launch (one phase line per call).  The environment variable
// STORM_RECORD_CHUNKS = K >= 1 fixes the count instead.
static int record_chunks(storm_mapper* m, const storm_reads_view* v, bool want_tb) {
#ifdef STORM_FUSED_PHASES
  (void)m;
  (void)v;
  (void)want_tb;
  return 1;
#else
  const int64_t n = v->n_reads;
  if (m->res->refs.size() != 1) return 1;
  if (const char* e = getenv("STORM_RECORD_CHUNKS")) {   // a fixed count (1 = one launch), for tests
    const int k = atoi(e);
    if (k >= 1) return (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)k, 8, n}));
  }
  if (n < 262144) return 1;
  int maxL = 0;
  for (int64_t i = 0; i < n; ++i) maxL = std::max(maxL, (int)v->len[i]);
  if (maxL <= 0 || maxL > MAXL) return 1;   // the one-launch path reports the bad length
  FusedTune t;
  if (!cached_tune(m, tune_key(m, maxL), t) || t.ms <= 0) return 1;
  const double tail_ms = t.ms;
  const double copy_ms = (double)n * (4.0 * sizeof(HitRec) + (want_tb ? (double)tb_ranks(m->p) * tb_width(m, maxL) : 0.0)) / 40e6;
  const int K = (int)std::floor(std::sqrt(copy_ms / tail_ms));
  return (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)K, 8, n / 32768}));
#endif
}

// A batched call with whole records for the caller, in K chunks of consecutive reads of the view:
// each chunk is one fused launch over its reads (by length within the chunk), and the hit map and
// tracebacks of chunk k - 1 go to the caller (DMA on the mapper's second stream, host threads
// copying out of the pinned halves) while chunk k runs on the device, so the copies of all but
// the last chunk leave the call's critical path.  A read's results do not depend on the other reads
// of its launch, so they are those of one launch over the view.
static int map_records_chunked(storm_mapper* m, const storm_reads_view* v, const RecordsOut& ro, int K,
                               storm_map_stats* st_out) {
  FusedState& F = m->fs;
  const int64_t n = v->n_reads;
  std::vector<int64_t> bounds((size_t)K + 1);
  for (int k = 0; k <= K; ++k) bounds[(size_t)k] = n * k / K;
  std::vector<int64_t>& order = m->order64;
  bool sorted_view = true;
  length_order_chunks(v, bounds, order, sorted_view, m->len_hist, &m->order32);
  int maxL = 0;
  for (int L = MAXL + 1; L > 0 && !maxL; --L)
    if (m->len_hist[(size_t)L]) maxL = L;
  const double t0 = now_s();
  const double ti = m->res->index_seconds;
  int rc = upload_reads(m, v, order, F.reads, &m->len_hist, &sorted_view, &m->order32);
  if (rc) return rc;
  ScoreDev sc{};
  fill_score(m->p, sc);
  const int64_t stride = (int64_t)maxL * dp_dir_words(sc.I) * 4;   // direction words per lane
  const int bpc = fused_blocks_per_cu(sc.W, sc.I, m->p.color);
  if (bpc <= 0) return fail(STORM_ERR_DEVICE, "fused kernel: no occupancy for this band class");
  const int max_grid = device_cus(m->device) * bpc;
  rc = ensure_fused_buffers(m, n, stride, max_grid);
  if (rc) return rc;
  HIPOK(launch_slots_init(F.slots.as<Slot>(), 4 * n, m->st));
  rc = ensure_ref_resident(m, 0);
  if (rc) return rc;
  FusedTune tune;
  rc = tune_fused(m, n, maxL, max_grid, tune);
  if (rc) return rc;
  std::vector<const void*> sig;
  hipGraphExec_t exec = nullptr;
  rc = fused_graph(m, 0, tune, sig, exec);
  if (rc) return rc;
  // the records kernel's device row -> caller row map: left on the device by the upload
  const bool identity = sorted_view;
  if (!identity && m->order_dev_n != n) return fail(STORM_ERR_DEVICE, "read order missing on the device");
  const int M = tb_ranks(m->p);
  const bool want_tb = m->want_tb && ro.tb;
  HIPOK(m->rec_dev.ensure((size_t)(4 * n) * sizeof(HitRec)));
  const int tw = tb_width(m);   // compact records, expanded into the caller's
  if (want_tb) HIPOK(m->tb_dev.ensure((size_t)(n * M) * tw));
  HIPOK(m->mapped_dev.ensure(16));
  HIPOK(hipMemsetAsync(m->mapped_dev.p, 0, 4, m->st));
  if (!m->st_out) HIPOK(hipStreamCreateWithFlags(&m->st_out, hipStreamNonBlocking));
  if (!m->rec_ev) HIPOK(hipEventCreateWithFlags(&m->rec_ev, hipEventDisableTiming));
  if (!m->ev0) HIPOK(hipEventCreate(&m->ev0));
  if (!m->ev1) HIPOK(hipEventCreate(&m->ev1));
  HIPOK(m->ranges.ensure((size_t)K * 8));
  int32_t* rg = (int32_t*)m->ranges.p;   // pinned: the range copies do not wait for the host
  // the caller's rows of chunk k (its device rows hold the same reads) leave on st_out once the
  // chunk's records kernel has run
  auto download = [&](int k) -> int {
    const int64_t a0 = bounds[(size_t)k], a1 = bounds[(size_t)k + 1];
    HIPOK(hipStreamWaitEvent(m->st_out, m->rec_ev, 0));
    std::vector<CopySeg> out = {{ro.hits + 4 * a0, m->rec_dev.as<HitRec>() + 4 * a0, (size_t)(4 * (a1 - a0)) * sizeof(HitRec)}};
    if (want_tb)
      out.push_back({ro.tb + a0 * M * TBB, m->tb_dev.as<uint8_t>() + a0 * M * tw, (size_t)((a1 - a0) * M) * tw, (size_t)tw,
                     (size_t)TBB});
    return staged_copy(m, out, true, m->st_out);
  };
  PassStats ps;
  unsigned long long totals[4] = {0, 0, 0, 0};
  int64_t n_fallback = 0;
  int pass_id = 0;
  std::vector<uint8_t>& ovf = m->ovf_host;
  if ((int64_t)ovf.size() < n) ovf.resize((size_t)n);
  for (int k = 0; k < K; ++k) {
    const int64_t r0 = bounds[(size_t)k], r1 = bounds[(size_t)k + 1];
    rg[2 * k] = (int32_t)r1;
    rg[2 * k + 1] = (int32_t)r0;
    HIPOK(hipMemcpyAsync(F.n_dev.p, rg + 2 * k, 8, hipMemcpyHostToDevice, m->st));
    HIPOK(hipEventRecord(m->ev0, m->st));
    HIPOK(hipGraphLaunch(exec, m->st));
    HIPOK(hipEventRecord(m->ev1, m->st));
    if (k > 0) {   // chunk k - 1's results out while chunk k runs
      rc = download(k - 1);
      if (rc) return rc;
    }
    unsigned long long st3[4];
    HIPOK(hipMemcpyAsync(st3, F.stats.p, sizeof st3, hipMemcpyDeviceToHost, m->st));
    HIPOK(hipMemcpyAsync(ovf.data() + r0, F.overflow.as<uint8_t>() + r0, (size_t)(r1 - r0), hipMemcpyDeviceToHost, m->st));
    HIPOK(hipStreamSynchronize(m->st));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, m->ev0, m->ev1);
    ps.ms_filter += ms;   // the fused kernel carries every stage
    for (int q = 0; q < 4; ++q) totals[q] += st3[q];
    std::vector<int64_t> over;
    for (int64_t i = r0; i < r1; ++i)
      if (ovf[(size_t)i]) over.push_back(i);
    n_fallback += (int64_t)over.size();
    if (!over.empty()) {
      rc = staged_subset(m, over, 0, pass_id, ps);
      if (rc) return rc;
    }
    HIPOK(launch_hit_records(F.slots.as<Slot>(), identity ? nullptr : m->order_dev.as<int32_t>(), r0, r1, M,
                             F.cells.as<uint8_t>(), m->rec_dev.as<HitRec>(), want_tb ? m->tb_dev.as<uint8_t>() : nullptr, tw,
                             m->mapped_dev.as<int>(), m->st));
    HIPOK(hipEventRecord(m->rec_ev, m->st));
  }
  int mapped_d = 0;
  HIPOK(hipMemcpyAsync(&mapped_d, m->mapped_dev.p, 4, hipMemcpyDeviceToHost, m->st));
  rc = download(K - 1);
  if (rc) return rc;
  HIPOK(hipStreamSynchronize(m->st));
  const int64_t mapped = mapped_d;
  m->finish_seconds = 0;
  return batched_stats(m, n, mapped, totals, ps, n_fallback, t0, ti, st_out);
}

static int map_view_batched(storm_mapper* m, const storm_reads_view* v, std::vector<HostSlot>& out_slots,
                            storm_map_stats* st_out, const HitsOut* ho, const RecordsOut* ro, const FinishFn* fin) {
  FusedState& F = m->fs;
  const int64_t n = v->n_reads;
  if (n <= 0) return fail(STORM_ERR_INVALID, "no reads");
  if (n > (int64_t)INT32_MAX / 4) return fail(STORM_ERR_INVALID, "batch too large");
  if (ro && !ho && !fin) {
    const int K = record_chunks(m, v, m->want_tb && ro->tb);
    if (K > 1) return map_records_chunked(m, v, *ro, K, st_out);
  }
  // device order: by length (tasks then hold reads of similar length)
  std::vector<int64_t>& order = m->order64;
  bool sorted_view = true;
  length_order_chunks(v, {0, n}, order, sorted_view, m->len_hist, &m->order32);
  int maxL = 0;
  for (int L = MAXL + 1; L > 0 && !maxL; --L)
    if (m->len_hist[(size_t)L]) maxL = L;
  const double t0 = now_s();
  const double ti = m->res->index_seconds;
  int rc = upload_reads(m, v, order, F.reads, &m->len_hist, &sorted_view, &m->order32);
  if (rc) return rc;
  PassStats ps;
  unsigned long long totals[4] = {0, 0, 0, 0};
  int64_t n_fallback = 0;
  FallbackFn fb = [&](const std::vector<int64_t>& over, int ref, int& pass_id, PassStats& pst) {
    return staged_subset(m, over, ref, pass_id, pst);
  };
  rc = fused_core(m, n, maxL, fb, ps, totals, n_fallback);
  if (rc) return rc;
  int64_t mapped = 0;
  const double tf = now_s();
  if (fin) rc = (*fin)(FinishCtx{F.slots.as<Slot>(), F.cells.as<uint8_t>(), &F.reads, &order, n}, mapped);
  else rc = finish_view(m, F.slots.as<Slot>(), F.cells.as<uint8_t>(), order, n, ho, ro, out_slots, mapped);
  if (rc) return rc;
  m->finish_seconds = fin ? now_s() - tf : 0;
  return batched_stats(m, n, mapped, totals, ps, n_fallback, t0, ti, st_out);
}
static int batched_stats(storm_mapper* m, int64_t n, int64_t mapped, const unsigned long long totals[4],
                         const PassStats& ps, int64_t n_fallback, double t0, double ti, storm_map_stats* st_out) {
  if (st_out) {
    st_out->n_reads = n;
    st_out->n_mapped = mapped;
    st_out->n_candidates = (int64_t)totals[0] + ps.cand;
    st_out->n_filter_pass = (int64_t)totals[1] + ps.rec;
    st_out->n_accepted = (int64_t)totals[2];
    st_out->n_hits = (int64_t)totals[3] + ps.hits;
    st_out->n_fallback = n_fallback;
    st_out->seconds_index = m->res->index_seconds - ti;
    st_out->seconds_map = now_s() - t0 - st_out->seconds_index - m->finish_seconds;
    st_out->gpu_ms_seed = ps.ms_seed;
    st_out->gpu_ms_filter = ps.ms_filter;
    st_out->gpu_ms_align = ps.ms_align;
    st_out->gpu_ms_replay = ps.ms_replay;
    st_out->gpu_ms_lookup = ps.ms_lookup;
  }
  return STORM_OK;
}

int load_resident(const char* genome_path, const storm_params& p, std::shared_ptr<Resident>& out) {
  std::shared_ptr<Resident> r = std::make_shared<Resident>();
  r->seeds_str = p.seeds ? p.seeds : default_seeds(p.color);
  std::string err;
  int rc = parse_seeds(r->seeds_str.c_str(), r->seeds, err);
  if (rc) return fail(rc, "%s", err.c_str());
  rc = load_reference(genome_path, p, r->refs, err);
  if (rc) return fail(rc, "%s", err.c_str());
  for (size_t k = 0; k < r->refs.size(); ++k) r->rstate.emplace_back(new RefState());
  out = std::move(r);
  return STORM_OK;
}

// index lookups of a call, from its length histogram (length_order_chunks)
static int64_t count_lookups(const storm_mapper* m, const storm_reads_view* v) {
  (void)v;
  const int strands = ((m->p.sense_mask & 1) ? 1 : 0) + ((m->p.sense_mask & 2) ? 1 : 0);
  int64_t n = 0;
  for (int L = 1; L <= MAXL && L < (int)m->len_hist.size(); ++L) {
    const int64_t c = m->len_hist[(size_t)L];
    if (!c) continue;
    for (const SeedSpec& s : m->res->seeds)
      n += c * (s.pos.empty() ? std::max<int64_t>(0, L - s.span + 1) : (int64_t)s.pos.size());
  }
  return n * strands * (int64_t)m->res->refs.size();
}

int map_view(storm_mapper* m, const storm_reads_view* v, int mode, std::vector<HostSlot>& out_slots,
             storm_map_stats* st_out, const HitsOut* ho, const RecordsOut* ro, const FinishFn* fin) {
  m->finish_seconds = 0;
  m->order_dev_n = -1;
  int rc = (mode == STORM_MODE_BATCHED) ? map_view_batched(m, v, out_slots, st_out, ho, ro, fin)
                                         : map_view_staged(m, v, mode, out_slots, st_out, ho, ro, fin);
  if (rc == STORM_OK && st_out) st_out->n_lookups = count_lookups(m, v);
  return rc;
}


}  // namespace storm


const char* storm::default_seeds(int color) {   // seed.h:29-32
  return color ? "####----##--##-####;####-###--#----####;####-####-####"
               : "####--#-#--##--####;#####--###-#-###;###-##-#######";
}

int storm::check_params(const storm_params* p) {
  if (!p) return fail(STORM_ERR_INVALID, "no parameters");
  if (p->allowed_indels < 0 || p->allowed_indels > MAX_I)
    return fail(STORM_ERR_INVALID, "-i %d outside 0..%d", p->allowed_indels, MAX_I);
  if (p->max_hits < 0 || p->max_hits > 4) return fail(STORM_ERR_INVALID, "-M must be in 1..4 (0: ordered mode)");
  if (p->max_hits == 0 && p->print_all)   // main.c:377-382
    return fail(STORM_ERR_INVALID, "-A (all reads in the SAM) needs the unordered mapping (-M <number>)");
  if (p->match <= 0 || p->mismatch > 0 || p->gap_open > 0 || p->gap_extend > 0)
    return fail(STORM_ERR_INVALID, "scores: match > 0, mismatch/gap penalties <= 0 (main.c:262-291)");
  if (p->sense_mask < 1 || p->sense_mask > 3) return fail(STORM_ERR_INVALID, "sense_mask must be 1, 2 or 3");
  return STORM_OK;
}

// =============================================================================== C ABI
using namespace storm;

extern "C" {

const char* storm_last_error(void) { return g_err.c_str(); }

void storm_params_default(storm_params* p, int color) {   // alignment.h:24-28,51-52; main.c:37-38
  memset(p, 0, sizeof(*p));
  p->color = color ? 1 : 0;
  p->match = 5; p->mismatch = -4; p->gap_open = -9; p->gap_extend = -4;
  p->allowed_indels = 3;
  p->t = 115; p->z = 100;
  p->qual_base = 33;
  p->stdev_distance = 100;
  p->sense_mask = 3;
  p->max_hits = 1;
  p->seeds = nullptr;
}

int storm_mapper_create(int device, const char* genome_path, const storm_params* params, storm_mapper** out) {
  *out = nullptr;
  int rc = check_params(params);
  if (rc) return rc;
  std::unique_ptr<storm_mapper> m(new storm_mapper());
  m->device = device;
  m->p = *params;
  rc = load_resident(genome_path, *params, m->res);
  if (rc) return rc;
  m->p.seeds = m->res->seeds_str.c_str();
  m->unmapped_str = params->unmapped_path ? params->unmapped_path : "";
  m->p.unmapped_path = params->unmapped_path ? m->unmapped_str.c_str() : nullptr;
  HIPOK(hipSetDevice(device));
  HIPOK(hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking));
  if (m->res->refs.size() == 1) {   // one sequence: keep genome + indexes resident in HBM
    rc = ensure_ref_resident(m.get(), 0);
    if (rc) return rc;
  }
  *out = m.release();
  return STORM_OK;
}

void storm_mapper_destroy(storm_mapper* m) {
  if (!m) return;
  (void)hipSetDevice(m->device);
  m->fs.drop_graphs();
  m->ps = PassScratch();
  m->res.reset();
  if (m->st) (void)hipStreamDestroy(m->st);
  delete m;
}

int64_t storm_mapper_free_bytes(storm_mapper* m) {
  size_t fr = 0, tot = 0;
  if (!m || hipSetDevice(m->device) != hipSuccess || hipMemGetInfo(&fr, &tot) != hipSuccess) return -1;
  return (int64_t)fr;
}

static int build_host_reads(const storm_reads_view* v, int color, HostReads& hr) {
  hr = HostReads();
  hr.n = v->n_reads;
  hr.db_name = v->db_name ? v->db_name : "";
  hr.code_stride = v->code_stride;
  hr.qual_stride = v->qual ? v->qual_stride : 0;
  hr.has_qual = v->qual != nullptr;
  hr.names.resize((size_t)hr.n);
  hr.len.assign(v->len, v->len + hr.n);
  hr.codes.assign(v->codes, v->codes + hr.n * v->code_stride);
  if (v->qual) hr.qual.assign(v->qual, v->qual + hr.n * v->qual_stride);
  if (color) {
    hr.first_base.assign(v->first_base ? v->first_base : nullptr, v->first_base ? v->first_base + hr.n : nullptr);
    hr.first_qual.assign(v->first_qual ? v->first_qual : nullptr, v->first_qual ? v->first_qual + hr.n : nullptr);
    if (hr.first_base.empty()) hr.first_base.assign((size_t)hr.n, 0);
    if (hr.first_qual.empty()) hr.first_qual.assign((size_t)hr.n, 0);
  }
  if (v->qual && v->no_qual) hr.no_qual.assign(v->no_qual, v->no_qual + hr.n);
  for (int64_t i = 0; i < hr.n; ++i) {
    hr.names[(size_t)i] = v->names ? std::string(v->names[i]) : ("r" + std::to_string(i));
    hr.max_len = std::max(hr.max_len, (int)v->len[i]);
  }
  return STORM_OK;
}

int storm_mapper_map(storm_mapper* m, const storm_reads_view* reads, int mode, const char* sam_path, int argc,
                     char** argv, storm_map_stats* stats) {
  if (!m || !reads || !sam_path) return fail(STORM_ERR_INVALID, "null argument");
  HIPOK(hipSetDevice(m->device));
  std::vector<HostSlot> slots;
  storm_map_stats local{};
  int rc = STORM_OK;
  if (m->p.max_hits == 0) {   // ordered mode: the genome map is built on the host from the hit map
    rc = map_view(m, reads, mode, slots, &local);
    if (rc) return rc;
    const double t0 = now_s();
    HostReads hr;
    build_host_reads(reads, m->p.color, hr);
    std::string err;
    rc = write_sam_ordered(sam_path, m->res->refs, hr, slots, m->p, argc, argv, err);
    if (rc) return fail(rc, "%s", err.c_str());
    if (m->p.unmapped_path) {   // run.c:1078-1086
      rc = write_unmapped_fastq(m->p.unmapped_path, hr, slots, m->p, err);
      if (rc) return fail(rc, "%s", err.c_str());
    }
    local.seconds_sam = now_s() - t0;
    if (stats) *stats = local;
    return STORM_OK;
  }
  // -M: the header on the host, the records formatted on the device straight from the hit map
  FILE* out = fopen(sam_path, "wb");
  if (!out) return fail(STORM_ERR_IO, "cannot write %s", sam_path);
  sam_header(out, m->res->refs, m->p.color != 0, reads->db_name ? reads->db_name : "", "unsorted", argc, argv);
  std::vector<uint8_t> unmapped;
  std::string names;
  std::vector<int64_t> name_off;
  const SamRows rows = sam_rows_of_view(reads, names, name_off);
  SamSink to_file = [&](const char* text, int64_t bytes, int64_t, int64_t, const int64_t*) {
    if (bytes && fwrite(text, 1, (size_t)bytes, out) != (size_t)bytes) return fail(STORM_ERR_IO, "SAM write failed");
    return STORM_OK;
  };
  FinishFn fin = [&](const FinishCtx& fc, int64_t& mapped) {
    int r = sam_device(m, rows, fc, to_file, false, mapped);
    if (r || !m->p.unmapped_path) return r;
    // -O: the reads with an empty hit map row (run.c:1078-1086)
    std::vector<storm_hit> hits((size_t)(4 * fc.n));
    RecordsOut ro{hits.data(), nullptr};
    std::vector<HostSlot> none;
    const bool keep = m->want_tb;
    m->want_tb = false;
    int64_t mp = 0;
    r = finish_view(m, fc.slots, fc.cells, *fc.order, fc.n, nullptr, &ro, none, mp);
    m->want_tb = keep;
    unmapped.resize((size_t)fc.n);
    for (int64_t i = 0; i < fc.n; ++i) unmapped[(size_t)i] = hits[(size_t)(4 * i)].ref_id == -1;
    return r;
  };
  rc = map_view(m, reads, mode, slots, &local, nullptr, nullptr, &fin);
  const bool write_ok = !ferror(out);
  if (fclose(out) != 0 || !write_ok) return rc ? rc : fail(STORM_ERR_IO, "error writing %s", sam_path);
  if (rc) return rc;
  local.seconds_sam = m->finish_seconds;
  if (m->p.unmapped_path) {
    const double t0 = now_s();
    HostReads hr;
    build_host_reads(reads, m->p.color, hr);
    std::vector<HostSlot> rows((size_t)(4 * hr.n), HostSlot{});
    for (int64_t i = 0; i < hr.n; ++i) rows[(size_t)(4 * i)].ref_id = unmapped[(size_t)i] ? -1 : 0;
    std::string err;
    rc = write_unmapped_fastq(m->p.unmapped_path, hr, rows, m->p, err);
    if (rc) return fail(rc, "%s", err.c_str());
    local.seconds_sam += now_s() - t0;
  }
  if (stats) *stats = local;
  return STORM_OK;
}

int storm_mapper_map_hits(storm_mapper* m, const storm_reads_view* reads, int mode, int32_t* score, int32_t* ref_id,
                          int32_t* ref_start, int32_t* sense, storm_map_stats* stats) {
  if (!m || !reads) return fail(STORM_ERR_INVALID, "null argument");
  HIPOK(hipSetDevice(m->device));
  std::vector<HostSlot> slots;
  storm_map_stats local{};
  HitsOut ho{score, ref_id, ref_start, sense};
  m->want_tb = false;
  int rc = map_view(m, reads, mode, slots, &local, &ho);
  m->want_tb = true;
  if (rc) return rc;
  if (stats) *stats = local;
  return STORM_OK;
}

int storm_mapper_map_records(storm_mapper* m, const storm_reads_view* reads, int mode, storm_hit* hits, uint8_t* tb,
                             storm_map_stats* stats) {
  if (!m || !reads || !hits) return fail(STORM_ERR_INVALID, "null argument");
  HIPOK(hipSetDevice(m->device));
  std::vector<HostSlot> slots;
  storm_map_stats local{};
  RecordsOut ro{hits, tb};
  m->want_tb = tb != nullptr;
  int rc = map_view(m, reads, mode, slots, &local, nullptr, &ro);
  m->want_tb = true;
  if (rc) return rc;
  if (stats) *stats = local;
  return STORM_OK;
}

int storm_map_files(int device, const char* genome_path, const char* reads_path, const char* qual_path,
                    const storm_params* params, int mode, const char* sam_path, int argc, char** argv,
                    storm_map_stats* stats) {
  int rc = check_params(params);
  if (rc) return rc;
  const double t0 = now_s();
  HostReads hr;
  bool on_device = false;
  if (!qual_path) {   // FASTQ: parsed, packed and ordered by tag on the device
    rc = load_reads_device(device, reads_path, *params, hr, on_device, nullptr);
    if (rc) return rc;
  }
  if (!on_device) {
    std::string err;
    rc = load_reads(reads_path, qual_path, *params, hr, err);
    if (rc) return fail(rc, "%s", err.c_str());
    sort_reads_by_tag(hr);   // run.c:826
  }
  const double t1 = now_s();
  storm_mapper* m = nullptr;
  rc = storm_mapper_create(device, genome_path, params, &m);
  if (rc) return rc;
  std::vector<const char*> names((size_t)hr.n);
  for (int64_t i = 0; i < hr.n; ++i) names[(size_t)i] = hr.names[(size_t)i].c_str();
  storm_reads_view v{};
  v.n_reads = hr.n;
  v.names = names.data();
  v.codes = hr.codes.data();
  v.code_stride = hr.code_stride;
  v.qual = hr.has_qual ? hr.qual.data() : nullptr;
  v.qual_stride = hr.qual_stride;
  v.len = hr.len.data();
  v.first_base = hr.first_base.empty() ? nullptr : hr.first_base.data();
  v.first_qual = hr.first_qual.empty() ? nullptr : hr.first_qual.data();
  v.no_qual = hr.no_qual.empty() ? nullptr : hr.no_qual.data();
  v.db_name = hr.db_name.c_str();
  storm_map_stats local{};
  rc = storm_mapper_map(m, &v, mode, sam_path, argc, argv, &local);
  local.seconds_index += m->res->index_seconds;
  storm_mapper_destroy(m);
  if (rc) return rc;
  local.seconds_load = t1 - t0;
  if (stats) *stats = local;
  return STORM_OK;
}

// ------------------------------------------------------------------------------ filter plan
struct storm_filter_plan {
  int device = 0;
  hipStream_t st = nullptr;
  DevGenome g;
  DevMem planes, len, cand_read, cand_pos, score, pass;
  FilterArgs a{};
  int W = 0;
  int64_t n = 0;
};

static void pack_read_planes(const uint8_t* codes, int L, uint32_t* out) {
  for (int w = 0; w < 2 * RW; ++w) out[w] = 0;
  for (int i = 0; i < L; ++i) {
    const int c = (codes[i >> 2] >> ((i & 3) << 1)) & 3;
    out[i >> 5] |= (uint32_t)(c & 1) << (i & 31);
    out[RW + (i >> 5)] |= (uint32_t)((c >> 1) & 1) << (i & 31);
  }
}

int storm_filter_plan_create(int device, const uint8_t* genome_packed, int64_t genome_codes, const uint8_t* reads_packed,
                             int64_t read_stride, const int32_t* read_len, int64_t n_reads, const int32_t* cand_read,
                             const int32_t* cand_pos, int64_t n_cand, int W, int match, int mismatch, int gap_open,
                             int gap_extend, int z, storm_filter_plan** out) {
  *out = nullptr;
  if (!(W == 2 || W == 4 || W == 8 || W == 16)) return fail(STORM_ERR_INVALID, "W must be 2, 4, 8 or 16");
  for (int64_t r = 0; r < n_reads; ++r)
    if (read_len[r] < W || read_len[r] > MAXL)
      return fail(STORM_ERR_INVALID, "read %lld length %d outside [W=%d, %d]", (long long)r, read_len[r], W, MAXL);
  for (int64_t c = 0; c < n_cand; ++c) {
    if (cand_read[c] < 0 || cand_read[c] >= n_reads) return fail(STORM_ERR_INVALID, "candidate %lld: bad read", (long long)c);
    const int64_t p = cand_pos[c];
    if (p < -(GPAD - 64) || p + read_len[cand_read[c]] + 2 * W > genome_codes + GPAD - 64)
      return fail(STORM_ERR_INVALID, "candidate %lld: window start %lld outside the padded genome", (long long)c,
                  (long long)p);
  }
  std::unique_ptr<storm_filter_plan> pl(new storm_filter_plan());
  pl->device = device;
  HIPOK(hipSetDevice(device));
  HIPOK(hipStreamCreateWithFlags(&pl->st, hipStreamNonBlocking));
  std::vector<uint8_t> codes((size_t)genome_codes);
  for (int64_t i = 0; i < genome_codes; ++i) codes[(size_t)i] = (genome_packed[i >> 2] >> ((i & 3) << 1)) & 3;
  int rc = upload_genome(codes.data(), nullptr, genome_codes, pl->g, pl->st);
  if (rc) return rc;
  std::vector<uint32_t> planes((size_t)n_reads * 2 * RW);
  for (int64_t r = 0; r < n_reads; ++r) pack_read_planes(reads_packed + r * read_stride, read_len[r], &planes[(size_t)r * 2 * RW]);
  HIPOK(pl->planes.ensure(planes.size() * 4 + 16));
  HIPOK(pl->len.ensure((size_t)n_reads * 4 + 16));
  HIPOK(pl->cand_read.ensure((size_t)n_cand * 4 + 16));
  HIPOK(pl->cand_pos.ensure((size_t)n_cand * 4 + 16));
  HIPOK(pl->score.ensure((size_t)n_cand + 16));
  HIPOK(pl->pass.ensure((size_t)n_cand + 16));
  HIPOK(hipMemcpyAsync(pl->planes.p, planes.data(), planes.size() * 4, hipMemcpyHostToDevice, pl->st));
  HIPOK(hipMemcpyAsync(pl->len.p, read_len, (size_t)n_reads * 4, hipMemcpyHostToDevice, pl->st));
  HIPOK(hipMemcpyAsync(pl->cand_read.p, cand_read, (size_t)n_cand * 4, hipMemcpyHostToDevice, pl->st));
  HIPOK(hipMemcpyAsync(pl->cand_pos.p, cand_pos, (size_t)n_cand * 4, hipMemcpyHostToDevice, pl->st));
  HIPOK(hipStreamSynchronize(pl->st));
  FilterArgs& a = pl->a;
  a.g = pl->g.view();
  a.rplanes = pl->planes.as<uint32_t>();
  a.rlen = pl->len.as<int32_t>();
  a.len_shift = 0;
  a.cand_read = pl->cand_read.as<int32_t>();
  a.cand_pos = pl->cand_pos.as<int32_t>();
  a.n = n_cand;
  a.score = pl->score.as<uint8_t>();
  a.pass = pl->pass.as<uint8_t>();
  a.mt = std::abs(match) & 0xff; a.mm = std::abs(mismatch) & 0xff;
  a.op = std::abs(gap_open) & 0xff; a.ex = std::abs(gap_extend) & 0xff;
  unsigned u = (unsigned)z;
  if (u >= 255) u = 254;
  a.thr = (int)u;
  pl->W = W;
  pl->n = n_cand;
  *out = pl.release();
  return STORM_OK;
}

int storm_filter_run(storm_filter_plan* pl, int reps, double* ms_per_launch) {
  if (!pl || reps < 1) return fail(STORM_ERR_INVALID, "bad plan or reps");
  HIPOK(hipSetDevice(pl->device));
  hipEvent_t a, b;
  HIPOK(hipEventCreate(&a));
  HIPOK(hipEventCreate(&b));
  HIPOK(hipEventRecord(a, pl->st));
  for (int k = 0; k < reps; ++k) HIPOK(launch_filter(pl->a, pl->W, pl->st));
  HIPOK(hipEventRecord(b, pl->st));
  HIPOK(hipEventSynchronize(b));
  float ms = 0;
  HIPOK(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (ms_per_launch) *ms_per_launch = ms / reps;
  return STORM_OK;
}

int storm_filter_fetch(storm_filter_plan* pl, uint8_t* scores, uint8_t* pass) {
  if (!pl) return fail(STORM_ERR_INVALID, "bad plan");
  HIPOK(hipSetDevice(pl->device));
  if (scores) HIPOK(hipMemcpyAsync(scores, pl->score.p, (size_t)pl->n, hipMemcpyDeviceToHost, pl->st));
  if (pass) HIPOK(hipMemcpyAsync(pass, pl->pass.p, (size_t)pl->n, hipMemcpyDeviceToHost, pl->st));
  HIPOK(hipStreamSynchronize(pl->st));
  return STORM_OK;
}

void storm_filter_plan_destroy(storm_filter_plan* pl) {
  if (!pl) return;
  (void)hipSetDevice(pl->device);
  if (pl->st) (void)hipStreamSynchronize(pl->st);
  pl->g = DevGenome();
  if (pl->st) (void)hipStreamDestroy(pl->st);
  delete pl;
}

// ------------------------------------------------------------------------------ per-call rows
// Process-global state like the reference filter's (alignment_simd.c:1032-1043): init sets the
// constants and read length; the genome registered once is resident in HBM.  The align calls are
// reentrant, as the reference's are (run.c:900 calls them from an OpenMP loop): each call takes a
// staging slot of its own from a pool.
namespace {
// One call's read planes, windows and pass flags, in pinned, coherent host memory that the
// kernel reads and writes directly: a call is one launch and one synchronisation, no copies.
struct CallStage {
  uint32_t planes[2 * RW];
  int32_t len;
  int32_t rd[64], ps[64];
  uint8_t pass[64];
};
struct CallSlot {
  CallStage* stage = nullptr;       // host address
  CallStage* stage_dev = nullptr;   // the same memory's device address
  hipStream_t st = nullptr;
};
struct CallState {
  DevGenome g;
  const unsigned char* code0 = nullptr;
  int64_t genome_codes = 0;
  int64_t lead_codes = 0;
  int W = 0, L = 0, thr = 0, mt = 5, mm = 4, op = 9, ex = 4;
  bool ready = false;
  std::mutex mu;                    // guards the slot pool
  std::vector<CallSlot> idle, all;
};
CallState g_call;
thread_local int t_call_status = STORM_OK;

int take_slot(CallSlot& out) {
  CallState& c = g_call;
  {
    std::lock_guard<std::mutex> lk(c.mu);
    if (!c.idle.empty()) {
      out = c.idle.back();
      c.idle.pop_back();
      return STORM_OK;
    }
  }
  CallSlot s;
  HIPOK(hipSetDevice(0));
  HIPOK(hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking));
  if (hipHostMalloc((void**)&s.stage, sizeof(CallStage), hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess ||
      hipHostGetDevicePointer((void**)&s.stage_dev, s.stage, 0) != hipSuccess) {
    if (s.stage) (void)hipHostFree(s.stage);
    (void)hipStreamDestroy(s.st);
    return fail(STORM_ERR_MEMORY, "alignment_gfx950: pinned staging memory");
  }
  std::lock_guard<std::mutex> lk(c.mu);
  c.all.push_back(s);
  out = s;
  return STORM_OK;
}

void give_slot(const CallSlot& s) {
  std::lock_guard<std::mutex> lk(g_call.mu);
  g_call.idle.push_back(s);
}
}  // namespace

int alignment_gfx950__set_genome(const unsigned char* sequence_alloc, size_t alloc_bytes, size_t code0_offset_bytes) {
  if (!sequence_alloc || code0_offset_bytes > alloc_bytes) return fail(STORM_ERR_INVALID, "bad genome buffer");
  CallState& c = g_call;
  HIPOK(hipSetDevice(0));
  hipStream_t st = nullptr;
  HIPOK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  // the whole allocation (padding included) becomes the device genome; code 0 sits at
  // code0_offset_bytes * 4 codes into it
  const int64_t codes = (int64_t)alloc_bytes * 4;
  std::vector<uint8_t> cv((size_t)codes);
  for (int64_t i = 0; i < codes; ++i) cv[(size_t)i] = (sequence_alloc[i >> 2] >> ((i & 3) << 1)) & 3;
  int rc = upload_genome(cv.data(), nullptr, codes, c.g, st);
  (void)hipStreamSynchronize(st);
  (void)hipStreamDestroy(st);
  if (rc) return rc;
  c.code0 = sequence_alloc + code0_offset_bytes;
  c.genome_codes = codes;
  c.lead_codes = (int64_t)code0_offset_bytes * 4;
  return STORM_OK;
}

static void call_init(int W, unsigned match, unsigned mismatch, unsigned gapopen, unsigned gapextends, unsigned threshold,
                      unsigned readlength) {
  CallState& c = g_call;
  c.W = W;
  c.L = (int)readlength;
  c.mt = (int)(match & 0xff); c.mm = (int)(mismatch & 0xff);
  c.op = (int)(gapopen & 0xff); c.ex = (int)(gapextends & 0xff);
  unsigned u = threshold;
  if (u >= 255) u = 254;   // alignment_simd.c:3643-3645
  c.thr = (int)u;
  c.ready = true;
}

void alignment_gfx950__init_octa(unsigned a, unsigned b, unsigned c, unsigned d, unsigned e, unsigned f) { call_init(2, a, b, c, d, e, f); }
void alignment_gfx950__init_quad(unsigned a, unsigned b, unsigned c, unsigned d, unsigned e, unsigned f) { call_init(4, a, b, c, d, e, f); }
void alignment_gfx950__init_pair(unsigned a, unsigned b, unsigned c, unsigned d, unsigned e, unsigned f) { call_init(8, a, b, c, d, e, f); }
void alignment_gfx950__init_mono(unsigned a, unsigned b, unsigned c, unsigned d, unsigned e, unsigned f) { call_init(16, a, b, c, d, e, f); }

// One call filters simd_mul windows of one read (run.c:660); returns bit k for window k.  On an
// error it returns 0 and the calling thread's alignment_gfx950__status() is the error code.
static unsigned call_align(int W, int mul, unsigned char* genome, int* pos_genome, unsigned char* read) {
  CallState& c = g_call;
  t_call_status = STORM_OK;
  if (!c.ready || c.W != W || !c.code0 || genome != c.code0) {
    t_call_status = fail(STORM_ERR_INVALID, "alignment_gfx950: init/set_genome missing or genome pointer not registered");
    return 0;
  }
  if (c.L < W || c.L > MAXL) { t_call_status = fail(STORM_ERR_INVALID, "read length %d not supported", c.L); return 0; }
  CallSlot slot;
  if (int rc = take_slot(slot)) { t_call_status = rc; return 0; }
  CallStage& in = *slot.stage;   // this call's own slot: no other call touches it until it is returned
  pack_read_planes(read, c.L, in.planes);
  in.len = c.L;
  for (int k = 0; k < mul; ++k) { in.rd[k] = 0; in.ps[k] = (int32_t)(pos_genome[k] + c.lead_codes); }
  FilterArgs a{};
  a.g = c.g.view();
  a.rplanes = slot.stage_dev->planes;
  a.rlen = &slot.stage_dev->len;
  a.len_shift = 0;
  a.cand_read = slot.stage_dev->rd;
  a.cand_pos = slot.stage_dev->ps;
  a.n = mul;
  a.pass = slot.stage_dev->pass;
  a.mt = c.mt; a.mm = c.mm; a.op = c.op; a.ex = c.ex; a.thr = c.thr;
  if (hipSetDevice(0) != hipSuccess || launch_filter(a, W, slot.st) != hipSuccess ||
      hipStreamSynchronize(slot.st) != hipSuccess) {
    t_call_status = fail(STORM_ERR_DEVICE, "alignment_gfx950: launch failed");
    give_slot(slot);
    return 0;
  }
  unsigned x = 0;
  for (int k = 0; k < mul; ++k) x |= (unsigned)(in.pass[k] & 1) << k;
  give_slot(slot);
  return x;
}

// simd_mul of the SSE2 row (run.c:109): octa 8, quad 4, pair 2, mono 1
unsigned alignment_gfx950__align_octa(unsigned char* g, int* p, unsigned char* r) { return call_align(2, 8, g, p, r); }
unsigned alignment_gfx950__align_quad(unsigned char* g, int* p, unsigned char* r) { return call_align(4, 4, g, p, r); }
unsigned alignment_gfx950__align_pair(unsigned char* g, int* p, unsigned char* r) { return call_align(8, 2, g, p, r); }
unsigned alignment_gfx950__align_mono(unsigned char* g, int* p, unsigned char* r) { return call_align(16, 1, g, p, r); }

void alignment_gfx950__clean(void) {
  CallState& c = g_call;
  std::lock_guard<std::mutex> lk(c.mu);
  (void)hipSetDevice(0);
  for (CallSlot& s : c.all) {
    (void)hipStreamSynchronize(s.st);
    (void)hipStreamDestroy(s.st);
    (void)hipHostFree(s.stage);
  }
  c.all.clear();
  c.idle.clear();
  c.g = DevGenome();
  c.code0 = nullptr;
  c.ready = false;
}

int alignment_gfx950__status(void) { return t_call_status; }

int alignment_gfx950__compatible_proc(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return 0;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 0;
  return strncmp(prop.gcnArchName, "gfx950", 6) == 0 ? 1 : 0;
}

}  // extern "C"
