// read_ingest.cpp -- FASTQ reads loaded on the device (load_reads_db + sort_reads_db,
// reads_data.c:59-67,351-566) and the C ABI storm_reads_* that exposes both loaders.
//
// The device path covers FASTQ files whose records are exactly four lines (header '@', sequence,
// '+' line, quality) after the leading '#' comment lines, which is what sequencers and the
// synthetic generators write.  FASTA and csfasta(+.qual) files, and FASTQ files with blank lines
// between records, go through the host loader (host_io.cpp), whose result is the same.
#include <cstring>
#include <thread>

#include "mapper.h"

namespace storm {

namespace {

bool is_fastq_path(const char* path) {
  const char* dot = strrchr(path, '.');
  return dot && (!strcmp(dot, ".fastq") || !strcmp(dot, ".fq") || !strcmp(dot, ".fastq2") || !strcmp(dot, ".fq2"));
}

bool read_all(const char* path, std::string& data) {
  FILE* f = fopen(path, "rb");
  if (!f) return false;
  fseek(f, 0, SEEK_END);
  long n = ftell(f);
  fseek(f, 0, SEEK_SET);
  data.resize(n > 0 ? (size_t)n : 0);
  size_t got = n > 0 ? fread(&data[0], 1, (size_t)n, f) : 0;
  fclose(f);
  return got == data.size();
}

// reads_data.c:460-474: the name runs from after the '>' / '@' to the first blank, '@' or
// non-printable character, at most 255 characters
std::string header_name(const char* line, size_t avail) {
  size_t k = 1;
  while (k < avail && (unsigned char)line[k] > ' ' && (unsigned char)line[k] <= '~' && line[k] != '@') {
    ++k;
    if (k >= 256) break;
  }
  return std::string(line + 1, k - 1);
}

// reads_data.c:538-564
std::string common_name_prefix(const std::string& a, const std::string& b) {
  size_t l = 0, t = 0;
  while (l < a.size() && l < b.size() && a[l] == b[l]) {
    unsigned char c = (unsigned char)a[l];
    if ((t == 0 && ((c >= '!' && c <= ')') || (c >= '+' && c <= '<') || (c >= '>' && c <= '~'))) ||
        (t > 0 && c >= '!' && c <= '~'))
      t = l + 1;
    ++l;
  }
  return a.substr(0, t);
}

template <class F>
void parallel_for(int64_t n, F f) {
  const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(16, std::min<int64_t>((int64_t)std::thread::hardware_concurrency(), n / 4096 + 1)));
  std::vector<std::thread> th;
  const int64_t per = (n + nt - 1) / nt;
  for (int t = 1; t < nt; ++t) th.emplace_back([&, t] { for (int64_t i = t * per; i < std::min(n, (t + 1) * per); ++i) f(i); });
  for (int64_t i = 0; i < std::min(n, per); ++i) f(i);
  for (auto& x : th) x.join();
}

}  // namespace

int load_reads_device(int device, const char* path, const storm_params& p, HostReads& out, bool& handled,
                      double* device_ms) {
  handled = false;
  if (!is_fastq_path(path)) return STORM_OK;
  std::string data;
  if (!read_all(path, data)) return fail(STORM_ERR_IO, "cannot read %s", path);
  // leading comment lines; "# Title: " names the read group (reads_data.c:380-390)
  size_t off = 0;
  std::string title;
  while (off < data.size()) {
    size_t e = data.find('\n', off);
    if (e == std::string::npos) e = data.size();
    size_t t = e;
    while (t > off && (data[t - 1] == '\r' || data[t - 1] == '\n')) --t;
    if (t == off || data[off] != '#') break;
    if (title.empty() && t - off >= 9 && data.compare(off, 9, "# Title: ") == 0) title = data.substr(off + 9, t - off - 9);
    off = e + 1;
  }
  if (off >= data.size()) return STORM_OK;
  const int64_t nb = (int64_t)(data.size() - off);
  HIPOK(hipSetDevice(device));
  hipStream_t st = nullptr;
  HIPOK(hipStreamCreate(&st));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } guard{st};
  hipEvent_t e0, e1;
  HIPOK(hipEventCreate(&e0));
  HIPOK(hipEventCreate(&e1));
  struct EventGuard {
    hipEvent_t a, b;
    ~EventGuard() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
  } eguard{e0, e1};
  DevMem text, counts, first, nl, tmp;
  HIPOK(text.ensure((size_t)nb + 16));
  HIPOK(hipMemcpyAsync(text.p, data.data() + off, (size_t)nb, hipMemcpyHostToDevice, st));
  HIPOK(hipEventRecord(e0, st));
  const int64_t nc = nl_chunks(nb);
  HIPOK(counts.ensure((size_t)(nc + 1) * 4));
  HIPOK(first.ensure((size_t)(nc + 1) * 4));
  HIPOK(hipMemsetAsync(counts.as<int32_t>() + nc, 0, 4, st));
  HIPOK(launch_nl_count(text.as<uint8_t>(), nb, counts.as<int32_t>(), nc, st));
  size_t tb = 0;
  HIPOK(scan_exclusive_i32(counts.as<int32_t>(), first.as<int32_t>(), nc + 1, nullptr, tb, st));
  HIPOK(tmp.ensure(tb));
  HIPOK(scan_exclusive_i32(counts.as<int32_t>(), first.as<int32_t>(), nc + 1, tmp.p, tb, st));
  int32_t n_nl32 = 0;
  HIPOK(hipMemcpyAsync(&n_nl32, first.as<int32_t>() + nc, 4, hipMemcpyDeviceToHost, st));
  HIPOK(hipStreamSynchronize(st));
  const int64_t n_nl = n_nl32;
  const int64_t n_lines = n_nl + (data.back() != '\n' ? 1 : 0);
  if (n_lines <= 0 || n_lines % 4 != 0) return STORM_OK;   // not four lines per record: host loader
  HIPOK(nl.ensure((size_t)(n_nl + 1) * 8));
  HIPOK(launch_nl_positions(text.as<uint8_t>(), nb, first.as<int32_t>(), nc, nl.as<int64_t>(), st));
  const int64_t n = n_lines / 4;
  DevMem len, header, hasq;
  HIPOK(len.ensure((size_t)n * 4));
  HIPOK(header.ensure((size_t)n * 8));
  HIPOK(hasq.ensure((size_t)n));
  IngestArgs a{};
  a.text = text.as<uint8_t>();
  a.n_bytes = nb;
  a.nl = nl.as<int64_t>();
  a.n_nl = n_nl;
  a.n_reads = n;
  a.color = p.color;
  a.qual_base = p.qual_base;
  a.len = len.as<int32_t>();
  a.header = header.as<int64_t>();
  a.hasq = hasq.as<uint8_t>();
  HIPOK(launch_fastq_len(a, st));
  std::vector<int32_t> hlen((size_t)n);
  std::vector<uint8_t> hq((size_t)n);
  HIPOK(hipMemcpyAsync(hlen.data(), a.len, (size_t)n * 4, hipMemcpyDeviceToHost, st));
  HIPOK(hipMemcpyAsync(hq.data(), a.hasq, (size_t)n, hipMemcpyDeviceToHost, st));
  HIPOK(hipStreamSynchronize(st));
  int maxL = 0, minL = 1 << 30;
  bool any_q = false;
  for (int64_t i = 0; i < n; ++i) {
    if (hlen[(size_t)i] == -1) return STORM_OK;   // irregular record: host loader
  }
  for (int64_t i = 0; i < n; ++i) {
    const int L = hlen[(size_t)i];
    if (L <= 0) { handled = true; return fail(STORM_ERR_INVALID, "read #%lld is empty", (long long)(i + 1)); }
    maxL = std::max(maxL, L);
    minL = std::min(minL, L);
    any_q |= hq[(size_t)i] != 0;
  }
  handled = true;
  if (p.strict_length && minL != maxL)
    return fail(STORM_ERR_INVALID, "reads have different lengths (the reference accepts one length per run)");
  if (maxL > 256) return fail(STORM_ERR_INVALID, "reads longer than 256 codes are not supported");
  const int64_t stride = ((maxL + 3) / 4 + 7) & ~7;
  DevMem codes, qual, tag, fb, fq, err, order0, order1, tag1, codes2, qual2, len2, fb2, fq2;
  HIPOK(codes.ensure((size_t)(n * stride)));
  HIPOK(hipMemsetAsync(codes.p, 0, (size_t)(n * stride), st));
  if (any_q) {
    HIPOK(qual.ensure((size_t)(n * stride)));
    HIPOK(hipMemsetAsync(qual.p, 0, (size_t)(n * stride), st));
  }
  HIPOK(tag.ensure((size_t)n * 8));
  HIPOK(fb.ensure((size_t)n + 16));
  HIPOK(fq.ensure((size_t)n + 16));
  HIPOK(hipMemsetAsync(fb.p, 0, (size_t)n, st));
  HIPOK(hipMemsetAsync(fq.p, 0, (size_t)n, st));
  HIPOK(err.ensure((size_t)n));
  a.codes = codes.as<uint8_t>();
  a.qual = any_q ? qual.as<uint8_t>() : nullptr;
  a.code_stride = stride;
  a.tag = tag.as<uint64_t>();
  a.first_base = fb.as<uint8_t>();
  a.first_qual = fq.as<uint8_t>();
  a.err = err.as<uint8_t>();
  HIPOK(launch_fastq_pack(a, st));
  // sort_reads_db (reads_data.c:59-67): stable order by tag
  HIPOK(order0.ensure((size_t)n * 4));
  HIPOK(order1.ensure((size_t)n * 4));
  HIPOK(tag1.ensure((size_t)n * 8));
  HIPOK(launch_iota_i32(order0.as<int32_t>(), n, st));
  int sel = 0;
  size_t sb = 0;
  HIPOK(sort_tags(tag.as<uint64_t>(), tag1.as<uint64_t>(), order0.as<int32_t>(), order1.as<int32_t>(), n, nullptr, sb,
                  sel, st));
  HIPOK(tmp.ensure(std::max(sb, tb)));
  HIPOK(sort_tags(tag.as<uint64_t>(), tag1.as<uint64_t>(), order0.as<int32_t>(), order1.as<int32_t>(), n, tmp.p, sb,
                  sel, st));
  const int32_t* ord = sel ? order1.as<int32_t>() : order0.as<int32_t>();
  const uint64_t* tags_sorted = sel ? tag1.as<uint64_t>() : tag.as<uint64_t>();
  HIPOK(codes2.ensure((size_t)(n * stride)));
  HIPOK(launch_permute_bytes(codes.as<uint8_t>(), stride, ord, n, codes2.as<uint8_t>(), st));
  if (any_q) {
    HIPOK(qual2.ensure((size_t)(n * stride)));
    HIPOK(launch_permute_bytes(qual.as<uint8_t>(), stride, ord, n, qual2.as<uint8_t>(), st));
  }
  HIPOK(len2.ensure((size_t)n * 4));
  HIPOK(launch_permute_i32(len.as<int32_t>(), ord, n, len2.as<int32_t>(), st));
  HIPOK(fb2.ensure((size_t)n + 16));
  HIPOK(fq2.ensure((size_t)n + 16));
  HIPOK(launch_permute_bytes(fb.as<uint8_t>(), 1, ord, n, fb2.as<uint8_t>(), st));
  HIPOK(launch_permute_bytes(fq.as<uint8_t>(), 1, ord, n, fq2.as<uint8_t>(), st));
  HIPOK(hipEventRecord(e1, st));
  std::vector<uint8_t> herr((size_t)n);
  std::vector<int32_t> hord((size_t)n);
  std::vector<int64_t> hhead((size_t)n);
  HIPOK(hipMemcpyAsync(herr.data(), a.err, (size_t)n, hipMemcpyDeviceToHost, st));
  HIPOK(hipMemcpyAsync(hord.data(), ord, (size_t)n * 4, hipMemcpyDeviceToHost, st));
  HIPOK(hipMemcpyAsync(hhead.data(), a.header, (size_t)n * 8, hipMemcpyDeviceToHost, st));
  out = HostReads();
  out.n = n;
  out.max_len = maxL;
  out.has_qual = any_q;
  out.code_stride = out.qual_stride = stride;
  out.codes.resize((size_t)(n * stride));
  out.len.resize((size_t)n);
  out.tag.resize((size_t)n);
  HIPOK(hipMemcpyAsync(out.codes.data(), codes2.p, (size_t)(n * stride), hipMemcpyDeviceToHost, st));
  HIPOK(hipMemcpyAsync(out.len.data(), len2.p, (size_t)n * 4, hipMemcpyDeviceToHost, st));
  HIPOK(hipMemcpyAsync(out.tag.data(), tags_sorted, (size_t)n * 8, hipMemcpyDeviceToHost, st));
  if (any_q) {
    out.qual.resize((size_t)(n * stride));
    HIPOK(hipMemcpyAsync(out.qual.data(), qual2.p, (size_t)(n * stride), hipMemcpyDeviceToHost, st));
  }
  if (p.color) {
    out.first_base.resize((size_t)n);
    out.first_qual.resize((size_t)n);
    HIPOK(hipMemcpyAsync(out.first_base.data(), fb2.p, (size_t)n, hipMemcpyDeviceToHost, st));
    HIPOK(hipMemcpyAsync(out.first_qual.data(), fq2.p, (size_t)n, hipMemcpyDeviceToHost, st));
  }
  HIPOK(hipStreamSynchronize(st));
  if (device_ms) {
    float ms = 0;
    HIPOK(hipEventElapsedTime(&ms, e0, e1));
    *device_ms = ms;
  }
  for (int64_t i = 0; i < n; ++i)
    if (herr[(size_t)i]) return fail(STORM_ERR_INVALID, "quality character below the -b base");
  // names in file order, then in tag order
  std::vector<std::string> names((size_t)n);
  const char* base = data.data() + off;
  parallel_for(n, [&](int64_t i) {
    const int64_t h = hhead[(size_t)i];
    names[(size_t)i] = header_name(base + h, (size_t)std::min<int64_t>(nb - h, 257));
  });
  out.db_name = title.empty() ? common_name_prefix(names.front(), names.back()) : title;
  if (any_q) {
    bool any_missing = false;
    for (int64_t i = 0; i < n; ++i) any_missing |= hq[(size_t)i] == 0;
    if (any_missing) {
      out.no_qual.resize((size_t)n);
      for (int64_t i = 0; i < n; ++i) out.no_qual[(size_t)i] = hq[(size_t)hord[(size_t)i]] == 0;
    }
  }
  out.names.resize((size_t)n);
  parallel_for(n, [&](int64_t i) { out.names[(size_t)i] = std::move(names[(size_t)hord[(size_t)i]]); });
  return STORM_OK;
}

}  // namespace storm

using namespace storm;

struct storm_reads {
  HostReads r;
  std::vector<const char*> names;
  double device_ms = 0;
  int on_device = 0;
};

extern "C" {

int storm_reads_load(int device, const char* reads_path, const char* qual_path, const storm_params* params,
                     int use_device, storm_reads** out) {
  if (!out || !reads_path || !params) return fail(STORM_ERR_INVALID, "storm_reads_load: null argument");
  *out = nullptr;
  int rc = check_params(params);
  if (rc) return rc;
  std::unique_ptr<storm_reads> h(new storm_reads());
  bool handled = false;
  if (use_device && !qual_path) {
    rc = load_reads_device(device, reads_path, *params, h->r, handled, &h->device_ms);
    if (rc) return rc;
  }
  if (!handled) {
    std::string err;
    rc = load_reads(reads_path, qual_path, *params, h->r, err);
    if (rc) return fail(rc, "%s", err.c_str());
    sort_reads_by_tag(h->r);   // run.c:826
  }
  h->on_device = handled ? 1 : 0;
  h->names.resize((size_t)h->r.n);
  for (int64_t i = 0; i < h->r.n; ++i) h->names[(size_t)i] = h->r.names[(size_t)i].c_str();
  *out = h.release();
  return STORM_OK;
}

int storm_reads_get(const storm_reads* rd, storm_reads_view* v, uint64_t* tags, int* loaded_on_device,
                    double* device_ms) {
  if (!rd || !v) return fail(STORM_ERR_INVALID, "storm_reads_get: null argument");
  const HostReads& r = rd->r;
  *v = storm_reads_view{};
  v->n_reads = r.n;
  v->names = rd->names.data();
  v->codes = r.codes.data();
  v->code_stride = r.code_stride;
  v->qual = r.has_qual ? r.qual.data() : nullptr;
  v->qual_stride = r.qual_stride;
  v->len = r.len.data();
  v->first_base = r.first_base.empty() ? nullptr : r.first_base.data();
  v->first_qual = r.first_qual.empty() ? nullptr : r.first_qual.data();
  v->db_name = r.db_name.c_str();
  v->no_qual = r.no_qual.empty() ? nullptr : r.no_qual.data();
  if (tags) memcpy(tags, r.tag.data(), r.tag.size() * 8);
  if (loaded_on_device) *loaded_on_device = rd->on_device;
  if (device_ms) *device_ms = rd->device_ms;
  return STORM_OK;
}

void storm_reads_destroy(storm_reads* rd) { delete rd; }

}  // extern "C"
