// server.cpp -- serving loop of the mapping path (storm_server_*): the genome and its seed
// indexes stay resident in HBM in both encodings (nucleotides and colors); requests of any size,
// read length, encoding and band (-i) are admitted while device memory is free and mapped by
// continuous batching.
//
// Design.  An admitted request owns a region of one device arena (its packed reads and its
// result slots), reserved against the HBM budget at admission and returned when its results
// have been copied out; a request waits in FIFO order while the head of the queue does not fit.
// `lanes` worker threads each own a stream and the buffers of one batch (a storm_mapper over the
// shared resident genome).  A lane that becomes free admits what fits, then takes the next batch:
// up to max_batch_reads reads, all of one parameter class (encoding, -i, -z, -t, scores, strands),
// drawn from every admitted request of that class that still has unbatched reads -- the tail of
// one request and the head of the next share a launch.  Requests are served in order of a
// deadline, submit time + 4 x their estimated remaining device time (read count x mean length^2 x
// filter lanes, priced by the measured time of earlier batches): a small request does not wait
// behind a large one, and a large one is only passed by requests that the same rule puts first,
// so it is never starved (or in admission order, STORM_SCHED_FIFO).  The batch's reads
// are gathered from their regions on the device, sorted by length, mapped by the persistent
// fused kernel (its overflow reads by the staged kernels), and the slots scattered back to each
// request's region.  Lanes run concurrently, so batches of different requests and classes
// overlap on the device, and host work of one lane (admission, uploads, result copies) hides
// behind the kernels of the others.  Each read's result depends only on the read, the genome and
// the parameters (-M mapping, hit_map.c:50-144), so it is the same whatever batch it lands in.
//
// Plain mode (the per-request baseline): one lane maps one whole request at a time with the
// staged kernels (one launch per stage and read length), as storm_mapper_map_hits does.
#include <algorithm>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <list>
#include <mutex>
#include <thread>

#include "mapper.h"

namespace storm {

// First-fit allocator over one device allocation.  Allocation and release are host bookkeeping
// only, so admitting or retiring a request never synchronises the device.
class Arena {
 public:
  static constexpr size_t ALIGN = 256;
  int init(size_t bytes) {
    bytes = bytes / ALIGN * ALIGN;
    if (bytes == 0) return fail(STORM_ERR_MEMORY, "serving arena: no device memory left for requests");
    HIPOK(mem_.ensure(bytes));
    cap_ = bytes;
    free_.clear();
    free_[0] = bytes;
    used_ = peak_ = 0;
    return STORM_OK;
  }
  static size_t round(size_t b) { return (b + ALIGN - 1) / ALIGN * ALIGN; }
  bool alloc(size_t bytes, size_t& off) {
    bytes = round(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      off = it->first;
      const size_t rest = it->second - bytes;
      free_.erase(it);
      if (rest) free_[off + bytes] = rest;
      used_ += bytes;
      peak_ = std::max(peak_, used_);
      return true;
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = round(bytes);
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  bool fits(size_t bytes) const {
    bytes = round(bytes);
    for (const auto& kv : free_)
      if (kv.second >= bytes) return true;
    return false;
  }
  uint8_t* base() const { return mem_.as<uint8_t>(); }
  size_t cap() const { return cap_; }
  size_t used() const { return used_; }
  size_t peak() const { return peak_; }

 private:
  DevMem mem_;
  std::map<size_t, size_t> free_;   // offset -> bytes
  size_t cap_ = 0, used_ = 0, peak_ = 0;
};

// parameters that change the kernels' work; requests batch together only within one class
struct ClassKey {
  int color, I, z, t, match, mismatch, gap_open, gap_extend, sense_mask;
  int max_hits, uniqueness, print_all, qual_base;   // SAM output options (one formatting pass per batch)
  bool operator==(const ClassKey& o) const { return memcmp(this, &o, sizeof o) == 0; }
};

static ClassKey class_of(const storm_params& p) {
  ClassKey k;
  memset(&k, 0, sizeof k);
  k.color = p.color; k.I = p.allowed_indels; k.z = p.z; k.t = p.t; k.match = p.match; k.mismatch = p.mismatch;
  k.gap_open = p.gap_open; k.gap_extend = p.gap_extend; k.sense_mask = p.sense_mask;
  k.max_hits = p.max_hits; k.uniqueness = p.uniqueness; k.print_all = p.print_all; k.qual_base = p.qual_base;
  return k;
}

struct Request {
  int64_t id = 0;
  storm_params p{};
  ClassKey cls{};
  int64_t n = 0, cs = 0, qs = 0;
  bool has_qual = false, has_fq = false;
  std::vector<uint8_t> codes, qual, fq;   // host copy (strides cs, qs)
  std::vector<int32_t> len;
  // device region: codes | qual | len | first_qual | hits, each 256-byte aligned
  size_t region_off = 0, region_bytes = 0;
  size_t o_codes = 0, o_qual = 0, o_len = 0, o_fq = 0, o_hits = 0;
  bool admitted = false, uploaded = false, waited = false;
  hipEvent_t ready = nullptr;             // recorded after the region's upload on the admitting lane
  int64_t next_row = 0;                   // reads not yet given to a batch
  int64_t rows_done = 0;
  double units_per_read = 0;              // work estimate: mean length^2 x filter lanes
  int batches = 0;
  double t_submit = 0, t_admit = 0, t_first = 0, t_done = 0;
  bool done = false;
  int status = STORM_OK;
  std::string err;
  std::vector<int32_t> hits;              // n * 4 * {score, ref_id, ref_start, sense}
  // SAM records (sam_records): the reads' names packed (empty: "r<row>") and no-quality flags;
  // each row's text as batches deliver it, then the request's text in row order
  std::string names;
  std::vector<int64_t> name_off;
  std::vector<uint8_t> no_qual;
  std::mutex sam_mu;
  std::string sam_parts;
  std::vector<int64_t> part_off;
  std::vector<int32_t> part_len;
  std::string sam;
  const char* name_of(int64_t row, int& len, char* buf) const {
    if (!name_off.empty()) {
      len = (int)(name_off[(size_t)row + 1] - name_off[(size_t)row]);
      return names.data() + name_off[(size_t)row];
    }
    len = snprintf(buf, 24, "r%lld", (long long)row);
    return buf;
  }
};

// a range of one request's reads inside a batch
struct Segment {
  Request* r;
  int64_t row0, row1;
};

}  // namespace storm

using namespace storm;

struct storm_server {
  int device = 0;
  storm_server_config cfg{};
  std::shared_ptr<Resident> res[2];   // [0] nucleotides, [1] colors
  Arena arena;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> pending;       // submitted, waiting for device memory (FIFO)
  std::list<Request*> active;         // admitted, not finished
  std::unordered_map<int64_t, std::unique_ptr<Request>> reqs;
  int64_t next_id = 1;
  bool stop = false;
  int busy = 0;
  storm_server_stats st{};
  double t0 = 0;
  // device seconds per work unit (a read of length L in a band of W filter lanes is L^2 W units:
  // its candidates grow with L and each costs L x 2W filter cells), from the batches run so far
  // (their wall time on a lane, so it includes the sharing of the device between lanes)
  double sec_per_unit = 4e-11;
  struct Lane {
    std::unique_ptr<storm_mapper> m;
    std::thread th;
    DevMem rows, regions;
    std::vector<ServeRow> h_rows;
    std::vector<ServeRegion> h_regions;
  };
  std::vector<std::unique_ptr<Lane>> lanes;
  SharedTune tune;   // the lanes share autotuner results
  // fault injection for the error-path tests: the upload of the k-th admitted region (1-based,
  // STORM_SERVER_FAIL_UPLOAD=k in the environment at creation) reports a failure without
  // touching the device; 0 = off
  int64_t fail_upload = 0;
  int64_t n_uploads = 0;   // under mu
};

namespace storm {

static size_t region_layout(Request& r) {
  size_t o = 0;
  auto take = [&](size_t bytes) { const size_t at = o; o += Arena::round(bytes); return at; };
  r.o_codes = take((size_t)(r.n * r.cs));
  r.o_qual = take(r.has_qual ? (size_t)(r.n * r.qs) : 0);
  r.o_len = take((size_t)r.n * 4);
  r.o_fq = take(r.has_fq ? (size_t)r.n : 0);
  r.o_hits = take((size_t)r.n * 16 * 4);
  return o;
}

static ServeRegion region_view(storm_server* s, const Request& r) {
  uint8_t* b = s->arena.base() + r.region_off;
  ServeRegion g{};
  g.codes = b + r.o_codes;
  g.qual = r.has_qual ? b + r.o_qual : nullptr;
  g.len = (const int32_t*)(b + r.o_len);
  g.first_qual = r.has_fq ? b + r.o_fq : nullptr;
  g.code_stride = (int32_t)r.cs;
  g.qual_stride = (int32_t)r.qs;
  g.hits = (int32_t*)(b + r.o_hits);
  return g;
}

static int upload_region(storm_server* s, storm_mapper* m, Request& r) {
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (++s->n_uploads == s->fail_upload) return fail(STORM_ERR_DEVICE, "injected upload failure (request %lld)", (long long)r.id);
  }
  uint8_t* b = s->arena.base() + r.region_off;
  HIPOK(hipMemcpyAsync(b + r.o_codes, r.codes.data(), r.codes.size(), hipMemcpyHostToDevice, m->st));
  if (r.has_qual) HIPOK(hipMemcpyAsync(b + r.o_qual, r.qual.data(), r.qual.size(), hipMemcpyHostToDevice, m->st));
  HIPOK(hipMemcpyAsync(b + r.o_len, r.len.data(), r.len.size() * 4, hipMemcpyHostToDevice, m->st));
  if (r.has_fq) HIPOK(hipMemcpyAsync(b + r.o_fq, r.fq.data(), r.fq.size(), hipMemcpyHostToDevice, m->st));
  HIPOK(hipEventRecord(r.ready, m->st));
  return STORM_OK;
}

static void select_class(storm_mapper* m, storm_server* s, const storm_params& p) {
  m->p = p;
  m->res = s->res[p.color ? 1 : 0];
  m->p.seeds = m->res->seeds_str.c_str();
  m->p.stdev_distance = s->cfg.stdev_distance;
  m->p.ignore_lowercase = s->cfg.ignore_lowercase;
}

// One batch of the continuous-batching path on lane `ln`.
static int run_batch(storm_server* s, storm_server::Lane& ln, const std::vector<Segment>& segs) {
  storm_mapper* m = ln.m.get();
  select_class(m, s, segs[0].r->p);
  // region table (one entry per request in the batch) and rows sorted by length
  std::vector<Request*> regs;
  ln.h_regions.clear();
  int64_t n = 0;
  for (const Segment& g : segs) {
    if (std::find(regs.begin(), regs.end(), g.r) == regs.end()) {
      regs.push_back(g.r);
      ln.h_regions.push_back(region_view(s, *g.r));
    }
    n += g.row1 - g.row0;
  }
  // the batch's reads ordered by length, longest first (counting sort; lengths are 1..MAXL,
  // checked at submit): the costliest tasks of the persistent kernel go first
  std::vector<int64_t> start(MAXL + 2, 0);
  int maxL = 0;
  for (const Segment& g : segs)
    for (int64_t row = g.row0; row < g.row1; ++row) start[(size_t)(MAXL - g.r->len[(size_t)row]) + 1]++;
  for (size_t b = 1; b < start.size(); ++b) start[b] += start[b - 1];
  ln.h_rows.resize((size_t)n);
  for (const Segment& g : segs) {
    const int32_t ri = (int32_t)(std::find(regs.begin(), regs.end(), g.r) - regs.begin());
    for (int64_t row = g.row0; row < g.row1; ++row) {
      const int L = g.r->len[(size_t)row];
      ln.h_rows[(size_t)start[(size_t)(MAXL - L)]++] = ServeRow{ri, (int32_t)row};
      maxL = std::max(maxL, L);
    }
  }
  for (Request* r : regs) HIPOK(hipStreamWaitEvent(m->st, r->ready, 0));
  HIPOK(ln.rows.ensure((size_t)n * sizeof(ServeRow)));
  HIPOK(ln.regions.ensure(ln.h_regions.size() * sizeof(ServeRegion)));
  HIPOK(hipMemcpyAsync(ln.rows.p, ln.h_rows.data(), (size_t)n * sizeof(ServeRow), hipMemcpyHostToDevice, m->st));
  HIPOK(hipMemcpyAsync(ln.regions.p, ln.h_regions.data(), ln.h_regions.size() * sizeof(ServeRegion),
                       hipMemcpyHostToDevice, m->st));
  DevReads& d = m->fs.reads;
  HIPOK(d.codes.ensure((size_t)n * SERVE_CS));
  HIPOK(d.qual.ensure((size_t)n * SERVE_CS));
  HIPOK(d.len.ensure((size_t)n * 4));
  HIPOK(d.first_qual.ensure((size_t)n));
  ServeGatherArgs ga{};
  ga.regions = ln.regions.as<ServeRegion>();
  ga.rows = ln.rows.as<ServeRow>();
  ga.n = n;
  ga.codes = d.codes.as<uint8_t>();
  ga.qual = d.qual.as<uint8_t>();
  ga.len = d.len.as<int32_t>();
  ga.first_qual = d.first_qual.as<uint8_t>();
  HIPOK(launch_serve_gather(ga, m->st));
  int rc = prep_device_reads(m, n, SERVE_CS, SERVE_CS, true, true, d);
  if (rc) return rc;
  PassStats ps;
  unsigned long long totals[4];
  int64_t n_fallback = 0;
  FallbackFn fb = [&](const std::vector<int64_t>& over, int ref, int& pass_id, PassStats& pst) {
    return staged_subset(m, over, ref, pass_id, pst);
  };
  rc = fused_core(m, n, maxL, fb, ps, totals, n_fallback);
  if (rc) return rc;
  HIPOK(launch_serve_scatter(m->fs.slots.as<Slot>(), ln.rows.as<ServeRow>(), n, ln.regions.as<ServeRegion>(), m->st));
  if (s->cfg.sam_records) {
    // the batch's SAM records (rows in batch order), each row's text handed to its request
    std::string names;
    std::vector<int64_t> name_off((size_t)n + 1);
    std::vector<uint8_t> no_qual((size_t)n);
    char buf[24];
    for (int64_t k = 0; k < n; ++k) {
      const ServeRow rw = ln.h_rows[(size_t)k];
      const Request* r = regs[(size_t)rw.region];
      int len = 0;
      const char* nm = r->name_of(rw.row, len, buf);
      name_off[(size_t)k] = (int64_t)names.size();
      names.append(nm, (size_t)len);
      no_qual[(size_t)k] = !r->has_qual || (!r->no_qual.empty() && r->no_qual[(size_t)rw.row]);
    }
    name_off[(size_t)n] = (int64_t)names.size();
    SamRows rows{};
    rows.n = n;
    rows.names = names.data();
    rows.name_off = name_off.data();
    rows.no_qual = no_qual.data();
    rows.has_qual = rows.has_first_qual = true;   // the gather writes levels (3 without quality) for every row
    rows.code_stride = rows.qual_stride = SERVE_CS;
    std::vector<int64_t> ident((size_t)n);
    for (int64_t k = 0; k < n; ++k) ident[(size_t)k] = k;
    const FinishCtx fc{m->fs.slots.as<Slot>(), m->fs.cells.as<uint8_t>(), &m->fs.reads, &ident, n};
    SamSink to_requests = [&](const char* text, int64_t, int64_t row0, int64_t c, const int64_t* off) {
      for (int64_t k = 0; k < c; ++k) {
        const ServeRow rw = ln.h_rows[(size_t)(row0 + k)];
        Request* r = regs[(size_t)rw.region];
        std::lock_guard<std::mutex> lk(r->sam_mu);
        r->part_off[(size_t)rw.row] = (int64_t)r->sam_parts.size();
        r->part_len[(size_t)rw.row] = (int32_t)(off[k + 1] - off[k]);
        r->sam_parts.append(text + (off[k] - off[0]), (size_t)(off[k + 1] - off[k]));
      }
      return STORM_OK;
    };
    int64_t mapped = 0;
    rc = sam_device(m, rows, fc, to_requests, true, mapped);
    if (rc) return rc;
  }
  HIPOK(hipStreamSynchronize(m->st));
  std::lock_guard<std::mutex> lk(s->mu);
  s->st.n_fallback += n_fallback;
  s->st.candidates += (int64_t)totals[0] + ps.cand;
  return STORM_OK;
}

// Plain mode: one whole request through the staged kernels from its host copy.
static int run_plain(storm_server* s, storm_server::Lane& ln, Request& r) {
  storm_mapper* m = ln.m.get();
  select_class(m, s, r.p);
  storm_reads_view v{};
  v.n_reads = r.n;
  v.codes = r.codes.data();
  v.code_stride = r.cs;
  v.qual = r.has_qual ? r.qual.data() : nullptr;
  v.qual_stride = r.qs;
  v.len = r.len.data();
  v.first_qual = r.has_fq ? r.fq.data() : nullptr;
  std::vector<HostSlot> slots;
  storm_map_stats ms{};
  const size_t ns = (size_t)r.n * 4;
  std::vector<int32_t> f((size_t)4 * ns);
  HitsOut ho{f.data(), f.data() + ns, f.data() + 2 * ns, f.data() + 3 * ns};
  int rc = STORM_OK;
  if (s->cfg.sam_records) {   // the request's SAM records in row order, then its hits
    std::vector<int64_t> name_off;
    std::string names;
    if (r.name_off.empty()) {
      name_off.resize((size_t)r.n + 1);
      char buf[24];
      for (int64_t k = 0; k < r.n; ++k) {
        int len = 0;
        const char* nm = r.name_of(k, len, buf);
        name_off[(size_t)k] = (int64_t)names.size();
        names.append(nm, (size_t)len);
      }
      name_off[(size_t)r.n] = (int64_t)names.size();
    }
    SamRows rows{};
    rows.n = r.n;
    rows.names = r.name_off.empty() ? names.data() : r.names.data();
    rows.name_off = r.name_off.empty() ? name_off.data() : r.name_off.data();
    rows.no_qual = r.no_qual.empty() ? nullptr : r.no_qual.data();
    rows.has_qual = r.has_qual;
    rows.has_first_qual = r.has_fq;
    rows.code_stride = r.cs;
    rows.qual_stride = r.qs;
    SamSink to_request = [&](const char* text, int64_t bytes, int64_t, int64_t, const int64_t*) {
      r.sam.append(text, (size_t)bytes);
      return STORM_OK;
    };
    FinishFn fin = [&](const FinishCtx& fc, int64_t& mapped) {
      int e = sam_device(m, rows, fc, to_request, false, mapped);
      if (e) return e;
      int64_t mp = 0;
      return finish_view(m, fc.slots, fc.cells, *fc.order, fc.n, &ho, nullptr, slots, mp);
    };
    m->want_tb = true;
    rc = map_view(m, &v, STORM_MODE_PLAIN, slots, &ms, nullptr, nullptr, &fin);
  } else {
    m->want_tb = false;
    rc = map_view(m, &v, STORM_MODE_PLAIN, slots, &ms, &ho);
  }
  if (rc) return rc;
  r.hits.resize(4 * ns);
  for (size_t k = 0; k < ns; ++k)
    for (size_t q = 0; q < 4; ++q) r.hits[4 * k + q] = f[q * ns + k];
  std::lock_guard<std::mutex> lk(s->mu);
  s->st.candidates += ms.n_candidates;
  return STORM_OK;
}

// under s->mu: requests whose regions fit are admitted in arrival order
static void admit(storm_server* s, std::vector<Request*>& to_upload) {
  while (!s->pending.empty()) {
    Request* r = s->pending.front();
    size_t off = 0;
    if (r->region_bytes > 0 && !s->arena.alloc(r->region_bytes, off)) {
      if (!r->waited) s->st.admission_waits++;
      r->waited = true;
      break;
    }
    r->region_off = off;
    r->admitted = true;
    r->t_admit = now_s();
    s->pending.pop_front();
    s->active.push_back(r);
    if (r->region_bytes > 0) to_upload.push_back(r);
    else r->uploaded = true;
  }
  s->st.arena_peak_bytes = (int64_t)s->arena.peak();
}

// under s->mu: a failed request (its upload failed, or a batch holding its reads failed) is never
// batched again: its unbatched reads are counted as done, and it goes to `retire` once no batch
// holds any of its reads
static void retire_failed(storm_server* s, std::vector<Request*>& retire) {
  for (Request* r : s->active) {
    if (r->status == STORM_OK || r->next_row >= r->n) continue;
    r->rows_done += r->n - r->next_row;
    r->next_row = r->n;
    if (r->rows_done == r->n) retire.push_back(r);
  }
}

// under s->mu: the next batch (empty when nothing is ready)
static void form_batch(storm_server* s, const std::vector<Request*>& own_uploads, std::vector<Segment>& segs) {
  segs.clear();
  auto ready = [&](Request* r) {
    return r->uploaded || std::find(own_uploads.begin(), own_uploads.end(), r) != own_uploads.end();
  };
  Request* head = nullptr;
  for (Request* r : s->active)
    if (r->next_row < r->n && r->status == STORM_OK && ready(r)) { head = r; break; }
  if (!head) return;
  const double now = now_s();
  if (s->cfg.mode == STORM_MODE_PLAIN) {   // one whole request
    segs.push_back(Segment{head, 0, head->n});
    head->next_row = head->n;
    head->t_first = now;
    return;
  }
  // the requests a batch may draw from, in the order it serves them: admission order (FIFO), or
  // by deadline = submit time + 4 x estimated remaining device time (STORM_SCHED_SIZE)
  std::vector<Request*> cand;
  for (Request* r : s->active)
    if (r->next_row < r->n && r->status == STORM_OK && ready(r)) cand.push_back(r);
  if (s->cfg.schedule == STORM_SCHED_SIZE) {
    auto deadline = [&](const Request* r) {
      return r->t_submit + 4.0 * (double)(r->n - r->next_row) * r->units_per_read * s->sec_per_unit;
    };
    std::stable_sort(cand.begin(), cand.end(),
                     [&](const Request* a, const Request* b) { return deadline(a) < deadline(b); });
    head = cand.front();
  }
  int64_t room = s->cfg.max_batch_reads;
  for (Request* r : cand) {
    if (room == 0) break;
    if (!(r->cls == head->cls)) continue;
    const int64_t k = std::min(room, r->n - r->next_row);
    segs.push_back(Segment{r, r->next_row, r->next_row + k});
    if (r->next_row == 0) r->t_first = now;
    r->next_row += k;
    room -= k;
    if (!s->cfg.cross_request) break;
  }
}

// under s->mu: the work units of a batch's segments
static double batch_units(const std::vector<Segment>& segs) {
  double u = 0;
  for (const Segment& g : segs) u += (double)(g.row1 - g.row0) * g.r->units_per_read;
  return u;
}

static bool has_work(storm_server* s) {
  if (!s->pending.empty() && (s->pending.front()->region_bytes == 0 || s->arena.fits(s->pending.front()->region_bytes)))
    return true;
  for (Request* r : s->active)
    if (r->next_row < r->n && r->uploaded) return true;   // failed ones included: they are retired
  return false;
}

// the request's results are ready on the device (all its batches synchronised): copy them out,
// retire its region, wake its waiter
static void complete(storm_server* s, storm_server::Lane& ln, Request* r) {
  int rc = STORM_OK;
  if (s->cfg.sam_records && s->cfg.mode != STORM_MODE_PLAIN && r->status == STORM_OK) {
    // the rows' records in row order (batches delivered them in their own order)
    std::lock_guard<std::mutex> lk(r->sam_mu);
    size_t total = 0;
    for (int64_t k = 0; k < r->n; ++k) total += (size_t)r->part_len[(size_t)k];
    r->sam.resize(total);
    size_t o = 0;
    for (int64_t k = 0; k < r->n; ++k) {
      const size_t l = (size_t)r->part_len[(size_t)k];
      memcpy(&r->sam[o], r->sam_parts.data() + r->part_off[(size_t)k], l);
      o += l;
    }
    std::string().swap(r->sam_parts);
  }
  if (s->cfg.mode != STORM_MODE_PLAIN && r->status == STORM_OK) {
    r->hits.resize((size_t)r->n * 16);
    const ServeRegion g = region_view(s, *r);
    hipError_t e = hipMemcpyAsync(r->hits.data(), g.hits, r->hits.size() * 4, hipMemcpyDeviceToHost, ln.m->st);
    if (e == hipSuccess) e = hipStreamSynchronize(ln.m->st);
    if (e != hipSuccess) rc = fail(STORM_ERR_DEVICE, "result copy: %s", hipGetErrorString(e));
  }
  std::lock_guard<std::mutex> lk(s->mu);
  if (rc != STORM_OK && r->status == STORM_OK) { r->status = rc; r->err = storm_last_error(); }
  if (r->region_bytes > 0) s->arena.release(r->region_off, r->region_bytes);
  s->active.remove(r);
  r->t_done = now_s();
  r->done = true;
  if (r->status == STORM_OK) s->st.completed++;
  else s->st.failed++;
  s->st.reads_done += r->n;
  s->cv_done.notify_all();
  s->cv_work.notify_all();
}

static void lane_loop(storm_server* s, storm_server::Lane* ln) {
  (void)hipSetDevice(s->device);
  std::vector<Request*> uploads, retire;
  std::vector<Segment> segs, run;
  for (;;) {
    retire.clear();
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_work.wait(lk, [&] { return s->stop || has_work(s); });
      if (s->stop) return;
      uploads.clear();
      admit(s, uploads);
      retire_failed(s, retire);
      form_batch(s, uploads, segs);
      if (!segs.empty()) {
        s->busy++;
        s->st.max_lanes_busy = std::max<int64_t>(s->st.max_lanes_busy, s->busy);
      }
    }
    for (Request* r : retire) complete(s, *ln, r);
    // uploads of the regions this lane admitted; other lanes wait on each region's event
    for (Request* r : uploads) {
      int rc = upload_region(s, ln->m.get(), *r);
      std::lock_guard<std::mutex> lk(s->mu);
      if (rc != STORM_OK) { r->status = rc; r->err = storm_last_error(); }
      r->uploaded = true;
      s->cv_work.notify_all();
    }
    if (segs.empty()) continue;
    // the segments of requests that failed meanwhile (an upload of this lane) are left out: their
    // regions hold no reads; they count as done and their requests complete as failed
    run.clear();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      for (const Segment& g : segs)
        if (g.r->status == STORM_OK) run.push_back(g);
    }
    const double tb = now_s();
    int rc = STORM_OK;
    if (!run.empty()) rc = (s->cfg.mode == STORM_MODE_PLAIN) ? run_plain(s, *ln, *run[0].r) : run_batch(s, *ln, run);
    std::string err = rc ? std::string(storm_last_error()) : std::string();
    std::vector<Request*> finished;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->busy--;
      s->st.batches++;
      const double dt = now_s() - tb, units = batch_units(run);
      s->st.gpu_busy_seconds += dt;
      if (rc == STORM_OK && units > 0 && s->cfg.mode != STORM_MODE_PLAIN)   // running price of a work unit
        s->sec_per_unit = 0.7 * s->sec_per_unit + 0.3 * (dt / units);
      int64_t nb = 0;
      std::vector<Request*> seen;
      for (Segment& g : segs) {
        nb += g.row1 - g.row0;
        if (std::find(seen.begin(), seen.end(), g.r) == seen.end()) seen.push_back(g.r);
        g.r->batches++;
        g.r->rows_done += g.row1 - g.row0;
        // a failed batch fails every request it held (their results were not computed)
        if (rc != STORM_OK && g.r->status == STORM_OK) { g.r->status = rc; g.r->err = err; }
        if (g.r->rows_done == g.r->n) finished.push_back(g.r);
      }
      s->st.batch_reads += nb;
      s->st.batch_requests += (int64_t)seen.size();
      s->st.max_requests_per_batch = std::max<int64_t>(s->st.max_requests_per_batch, (int64_t)seen.size());
    }
    for (Request* r : finished) complete(s, *ln, r);
  }
}

}  // namespace storm

static int wait_request(storm_server* s, int64_t id, int32_t* score, int32_t* ref_id, int32_t* ref_start, int32_t* sense,
                        storm_request_result* out, std::string* sam);

extern "C" {

void storm_server_config_default(storm_server_config* c) {
  memset(c, 0, sizeof *c);
  c->mode = STORM_MODE_BATCHED;
  c->lanes = 3;
  c->cross_request = 1;
  c->max_batch_reads = 1 << 18;
  c->hbm_budget_bytes = 0;
  c->hbm_reserve_bytes = 8LL << 30;
  c->encodings = 3;
  c->stdev_distance = 100;
  c->sam_records = 1;
  c->schedule = STORM_SCHED_SIZE;
}

int storm_server_create(int device, const char* genome_path, const storm_server_config* cfg, storm_server** out) {
  *out = nullptr;
  if (!genome_path || !cfg) return fail(STORM_ERR_INVALID, "null argument");
  if (cfg->lanes < 1 || cfg->lanes > 16) return fail(STORM_ERR_INVALID, "lanes must be in 1..16");
  if (cfg->max_batch_reads < 1 || cfg->max_batch_reads > (1LL << 24))
    return fail(STORM_ERR_INVALID, "max_batch_reads must be in 1..2^24");
  if (cfg->mode != STORM_MODE_PLAIN && cfg->mode != STORM_MODE_BATCHED) return fail(STORM_ERR_INVALID, "bad mode");
  if (cfg->encodings < 1 || cfg->encodings > 3) return fail(STORM_ERR_INVALID, "encodings must be 1, 2 or 3");
  if (cfg->schedule != STORM_SCHED_FIFO && cfg->schedule != STORM_SCHED_SIZE)
    return fail(STORM_ERR_INVALID, "schedule must be STORM_SCHED_FIFO or STORM_SCHED_SIZE");
  std::unique_ptr<storm_server> s(new storm_server());
  s->device = device;
  s->cfg = *cfg;
  if (s->cfg.mode == STORM_MODE_PLAIN) s->cfg.lanes = 1;
  HIPOK(hipSetDevice(device));
  for (int c = 0; c < 2; ++c) {
    if (!((cfg->encodings >> c) & 1)) continue;
    storm_params p;
    storm_params_default(&p, c);
    p.stdev_distance = cfg->stdev_distance;
    p.ignore_lowercase = cfg->ignore_lowercase;
    const char* sd = c ? cfg->seeds_color : cfg->seeds_nucleotide;
    p.seeds = sd;
    int rc = load_resident(genome_path, p, s->res[c]);
    if (rc) return rc;
    s->res[c]->pinned = true;
  }
  for (int k = 0; k < s->cfg.lanes; ++k) {
    std::unique_ptr<storm_server::Lane> ln(new storm_server::Lane());
    ln->m.reset(new storm_mapper());
    ln->m->device = device;
    ln->m->want_tb = s->cfg.sam_records != 0;   // tracebacks only for SAM records
    ln->m->shared_tune = &s->tune;
    HIPOK(hipStreamCreateWithFlags(&ln->m->st, hipStreamNonBlocking));
    s->lanes.push_back(std::move(ln));
  }
  // every sequence of both encodings resident, built once on the first lane's stream
  for (int c = 0; c < 2; ++c) {
    if (!s->res[c]) continue;
    storm_params p;
    storm_params_default(&p, c);
    p.stdev_distance = cfg->stdev_distance;
    storm_mapper* m = s->lanes[0]->m.get();
    m->p = p;
    m->res = s->res[c];
    for (size_t r = 0; r < s->res[c]->refs.size(); ++r) {
      int rc = ensure_ref_resident(m, r);
      if (rc) return rc;
    }
    m->res.reset();
  }
  if (s->cfg.mode != STORM_MODE_PLAIN) {
    size_t fr = 0, tot = 0;
    HIPOK(hipMemGetInfo(&fr, &tot));
    const int64_t avail = (int64_t)fr - s->cfg.hbm_reserve_bytes;
    // default: a quarter of the free HBM, the rest left to the lanes' batch buffers, the runtime's
    // per-queue scratch and other users of the device
    int64_t budget = s->cfg.hbm_budget_bytes > 0 ? std::min<int64_t>(s->cfg.hbm_budget_bytes, avail)
                                                 : std::min<int64_t>(avail, (int64_t)(fr / 4));
    if (budget <= 0) return fail(STORM_ERR_MEMORY, "free HBM (%lld B) is below the reserve (%lld B)", (long long)fr,
                                 (long long)s->cfg.hbm_reserve_bytes);
    int rc = s->arena.init((size_t)budget);
    if (rc) return rc;
    s->st.arena_bytes = (int64_t)s->arena.cap();
  }
  if (const char* f = getenv("STORM_SERVER_FAIL_UPLOAD")) s->fail_upload = atoll(f);
  s->t0 = now_s();
  storm_server* sp = s.get();
  for (auto& ln : s->lanes) ln->th = std::thread(lane_loop, sp, ln.get());
  *out = s.release();
  return STORM_OK;
}

int storm_server_submit(storm_server* s, const storm_reads_view* v, const storm_params* params, int64_t* id) {
  if (!s || !v || !params || !id) return fail(STORM_ERR_INVALID, "null argument");
  int rc = check_params(params);
  if (rc) return rc;
  if (params->max_hits == 0)
    return fail(STORM_ERR_INVALID, "the ordered mode (-M 0) places the reads of a whole run together; a request "
                                   "needs -M 1..4");
  const int color = params->color ? 1 : 0;
  if (!s->res[color]) return fail(STORM_ERR_INVALID, "the server holds no %s index", color ? "color" : "nucleotide");
  if (v->n_reads <= 0) return fail(STORM_ERR_INVALID, "no reads");
  if (v->n_reads > (int64_t)INT32_MAX / 16) return fail(STORM_ERR_INVALID, "request too large");
  std::unique_ptr<Request> r(new Request());
  r->p = *params;
  r->p.seeds = nullptr;
  r->p.unmapped_path = nullptr;   // a request returns hits, it writes no files
  r->cls = class_of(r->p);
  r->n = v->n_reads;
  r->cs = v->code_stride;
  r->qs = v->qual ? v->qual_stride : 0;
  r->has_qual = v->qual != nullptr;
  r->has_fq = color && v->first_qual != nullptr;
  const int wl = filter_lanes(params->allowed_indels >= 16 ? 15 : params->allowed_indels);
  for (int64_t i = 0; i < r->n; ++i) {
    const int L = v->len[i];
    if (L <= 0 || L > MAXL) return fail(STORM_ERR_INVALID, "read %lld has length %d (1..%d supported)", (long long)i, L, MAXL);
    if (L < wl) return fail(STORM_ERR_INVALID, "read %lld is shorter (%d) than the filter band lanes (%d)", (long long)i, L, wl);
    if (r->cs * 4 < L) return fail(STORM_ERR_INVALID, "code_stride %lld too small for length %d", (long long)r->cs, L);
    r->units_per_read += (double)L * L;
  }
  r->units_per_read *= (double)wl / (double)r->n;
  if (s->cfg.sam_records) {
    if (v->names) {
      r->name_off.resize((size_t)r->n + 1);
      for (int64_t i = 0; i < r->n; ++i) {
        r->name_off[(size_t)i] = (int64_t)r->names.size();
        r->names += v->names[i];
      }
      r->name_off[(size_t)r->n] = (int64_t)r->names.size();
    }
    if (v->qual && v->no_qual) r->no_qual.assign(v->no_qual, v->no_qual + r->n);
    if (s->cfg.mode != STORM_MODE_PLAIN) {
      r->part_off.assign((size_t)r->n, 0);
      r->part_len.assign((size_t)r->n, 0);
    }
  }
  r->codes.assign(v->codes, v->codes + r->n * r->cs);
  if (r->has_qual) r->qual.assign(v->qual, v->qual + r->n * r->qs);
  r->len.assign(v->len, v->len + r->n);
  if (r->has_fq) r->fq.assign(v->first_qual, v->first_qual + r->n);
  if (s->cfg.mode != STORM_MODE_PLAIN) {
    r->region_bytes = region_layout(*r);
    if (r->region_bytes > s->arena.cap())
      return fail(STORM_ERR_MEMORY, "request needs %zu device bytes, the serving budget is %zu", r->region_bytes,
                  s->arena.cap());
  }
  HIPOK(hipSetDevice(s->device));
  HIPOK(hipEventCreateWithFlags(&r->ready, hipEventDisableTiming));
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop) return fail(STORM_ERR_INVALID, "server is shutting down");
  r->id = s->next_id++;
  r->t_submit = now_s();
  *id = r->id;
  s->pending.push_back(r.get());
  s->st.submitted++;
  s->reqs[r->id] = std::move(r);
  s->cv_work.notify_all();
  return STORM_OK;
}

int storm_server_wait(storm_server* s, int64_t id, int32_t* score, int32_t* ref_id, int32_t* ref_start, int32_t* sense,
                      storm_request_result* out) {
  return wait_request(s, id, score, ref_id, ref_start, sense, out, nullptr);
}

}  // extern "C"

static int wait_request(storm_server* s, int64_t id, int32_t* score, int32_t* ref_id, int32_t* ref_start, int32_t* sense,
                        storm_request_result* out, std::string* sam) {
  if (!s) return fail(STORM_ERR_INVALID, "null server");
  std::unique_ptr<Request> r;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(STORM_ERR_INVALID, "unknown request %lld", (long long)id);
    Request* rp = it->second.get();
    s->cv_done.wait(lk, [&] { return rp->done; });
    r = std::move(it->second);
    s->reqs.erase(it);
  }
  if (r->ready) (void)hipEventDestroy(r->ready);
  if (out) {
    memset(out, 0, sizeof *out);
    out->n_reads = r->n;
    out->status = r->status;
    out->batches = r->batches;
    out->seconds_queued = r->t_admit - r->t_submit;
    out->seconds_latency = r->t_done - r->t_submit;
    out->t_submit = r->t_submit - s->t0;
    out->t_done = r->t_done - s->t0;
  }
  if (r->status != STORM_OK) return fail(r->status, "request %lld: %s", (long long)id, r->err.c_str());
  int64_t mapped = 0;
  for (int64_t i = 0; i < r->n; ++i) {
    for (int k = 0; k < 4; ++k) {
      const int32_t* h = &r->hits[(size_t)((i * 4 + k) * 4)];
      if (score) score[i * 4 + k] = h[0];
      if (ref_id) ref_id[i * 4 + k] = h[1];
      if (ref_start) ref_start[i * 4 + k] = h[2];
      if (sense) sense[i * 4 + k] = h[3];
    }
    if (r->hits[(size_t)(i * 16 + 1)] != -1) mapped++;
  }
  if (out) out->n_mapped = mapped;
  if (sam) sam->swap(r->sam);
  return STORM_OK;
}

extern "C" {

int storm_server_wait_sam(storm_server* s, int64_t id, char** sam, int64_t* sam_bytes, int32_t* score, int32_t* ref_id,
                          int32_t* ref_start, int32_t* sense, storm_request_result* out) {
  if (!s || !sam || !sam_bytes) return fail(STORM_ERR_INVALID, "null argument");
  *sam = nullptr;
  *sam_bytes = 0;
  if (!s->cfg.sam_records) return fail(STORM_ERR_INVALID, "the server was created without SAM records (sam_records = 0)");
  std::string text;
  int rc = wait_request(s, id, score, ref_id, ref_start, sense, out, &text);
  if (rc) return rc;
  char* buf = (char*)malloc(text.size() + 1);
  if (!buf) return fail(STORM_ERR_MEMORY, "SAM text of request %lld", (long long)id);
  memcpy(buf, text.data(), text.size());
  buf[text.size()] = 0;
  *sam = buf;
  *sam_bytes = (int64_t)text.size();
  return STORM_OK;
}

void storm_free(void* p) { free(p); }

int storm_server_get_stats(storm_server* s, storm_server_stats* out) {
  if (!s || !out) return fail(STORM_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->st;
  out->arena_used_bytes = (int64_t)s->arena.used();
  out->pending = (int64_t)s->pending.size();
  out->active = (int64_t)s->active.size();
  return STORM_OK;
}

void storm_server_destroy(storm_server* s) {
  if (!s) return;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    // finish what was submitted, then stop the lanes
    s->cv_done.wait(lk, [&] {
      for (auto& kv : s->reqs) if (!kv.second->done) return false;
      return true;
    });
    s->stop = true;
    s->cv_work.notify_all();
  }
  for (auto& ln : s->lanes) if (ln->th.joinable()) ln->th.join();
  (void)hipSetDevice(s->device);
  for (auto& kv : s->reqs) if (kv.second->ready) (void)hipEventDestroy(kv.second->ready);
  for (auto& ln : s->lanes) {
    ln->m->fs.drop_graphs();
    ln->m->res.reset();
    if (ln->m->st) (void)hipStreamSynchronize(ln->m->st);
  }
  std::vector<hipStream_t> streams;
  for (auto& ln : s->lanes) streams.push_back(ln->m->st);
  s->lanes.clear();
  for (hipStream_t q : streams) if (q) (void)hipStreamDestroy(q);
  delete s;
}

}  // extern "C"
