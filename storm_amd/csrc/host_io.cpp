// host_io.cpp -- host side of the mapping path: file loaders, seed parsing, CSR index build and
// the -M SAM writer.  These keep the reference's input conventions and output format so that the
// GPU path is a drop-in for its files (cited per function).
#include "host_io.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>
#include <thread>

namespace storm {

// ---------------------------------------------------------------------------------------------
// glibc rand() with its default seed: the reference draws rand() & 3 for each non-ACGT genome
// symbol in nucleotide mode (reference_data.c:119).  This is the published TYPE_3 additive
// feedback generator of glibc (r[i] = r[i-31] + r[i-3], output r >> 1, after 310 discarded
// steps) so the product does not depend on the process-wide rand() state.
class GlibcRand {
 public:
  explicit GlibcRand(uint32_t seed = 1) {
    int32_t r[34];
    r[0] = (int32_t)seed;
    for (int i = 1; i < 31; ++i) {
      int64_t v = (16807LL * r[i - 1]) % 2147483647LL;
      if (v < 0) v += 2147483647LL;
      r[i] = (int32_t)v;
    }
    for (int i = 31; i < 34; ++i) r[i] = r[i - 31];
    for (int i = 0; i < 34; ++i) ring_[i] = (uint32_t)r[i];
    pos_ = 34;
    for (int k = 34; k < 344; ++k) step();
  }
  int next() { return (int)(step() >> 1); }

 private:
  uint32_t step() {
    uint32_t v = ring_[(pos_ - 31) % 34] + ring_[(pos_ - 3) % 34];
    ring_[pos_ % 34] = v;
    ++pos_;
    return v;
  }
  uint32_t ring_[34];
  int64_t pos_;
};

uint32_t glibc_rand_next(uint32_t* state) {
  // state[0] holds a pointer-sized index into a lazily created generator table; kept for tests
  static thread_local std::unique_ptr<GlibcRand> g;
  if (state[0] == 0) { g.reset(new GlibcRand(1)); state[0] = 1; }
  return (uint32_t)g->next();
}

// ------------------------------------------------------------------------------------ seeds
// seed.c:25-135: "#-##" / "1011" patterns separated by ';', optional ":p1 p2" position list
int parse_seeds(const char* src, std::vector<SeedSpec>& out, std::string& err) {
  out.clear();
  if (!src || !*src) { err = "no seed given"; return STORM_ERR_INVALID; }
  std::string all(src);
  size_t start = 0;
  while (start <= all.size()) {
    size_t semi = all.find(';', start);
    std::string tok = all.substr(start, semi == std::string::npos ? std::string::npos : semi - start);
    start = (semi == std::string::npos) ? all.size() + 1 : semi + 1;
    if (tok.empty()) continue;
    SeedSpec s;
    size_t colon = tok.find(':');
    std::string pat = tok.substr(0, colon);
    s.span = (int)pat.size();
    for (int i = 0; i < s.span; ++i) {
      char c = pat[i];
      if (c == '#' || c == '1') s.offs.push_back(i);
      else if (c != '-' && c != '0') { err = "invalid seed pattern \"" + tok + "\""; return STORM_ERR_INVALID; }
    }
    s.weight = (int)s.offs.size();
    if (s.weight == 0) { err = "seed without '#': \"" + tok + "\""; return STORM_ERR_INVALID; }
    if (s.weight > 14) { err = "seed weight above 14 is not supported"; return STORM_ERR_INVALID; }
    if (s.span > 32) { err = "seed span above 32 is not supported"; return STORM_ERR_INVALID; }
    if (colon != std::string::npos && colon + 1 < tok.size()) {
      const char* q = tok.c_str() + colon + 1;
      while (*q) {
        while (*q == ' ' || *q == ',') ++q;
        if (!*q) break;
        char* end = nullptr;
        long v = strtol(q, &end, 10);
        if (end == q) break;
        s.pos.push_back((int)v);
        q = end;
      }
      if (s.pos.size() > 64) { err = "more than 64 explicit seed positions"; return STORM_ERR_INVALID; }
    }
    out.push_back(s);
  }
  if (out.empty()) { err = "no seed given"; return STORM_ERR_INVALID; }
  if (out.size() > 8) { err = "more than 8 seeds"; return STORM_ERR_INVALID; }
  return STORM_OK;
}

// ------------------------------------------------------------------------------------ files
static bool read_file(const char* path, std::string& data) {
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

// split into lines without the trailing "\r\n" characters
static void split_lines(const std::string& data, std::vector<std::pair<size_t, size_t>>& lines) {
  lines.clear();
  size_t s = 0, n = data.size();
  while (s < n) {
    size_t e = data.find('\n', s);
    if (e == std::string::npos) e = n;
    size_t t = e;
    while (t > s && (data[t - 1] == '\r' || data[t - 1] == '\n')) --t;
    lines.emplace_back(s, t - s);
    s = e + 1;
  }
}

static inline uint8_t base_symbol_code(int c) {   // codes.h:39
  return (uint8_t)((((c & 0x4) >> 1)) | (((c & 0x4) >> 2) ^ ((c & 0x2) >> 1)));
}

// reference_data.c:14-252: multi-FASTA, whitespace skipped, upper-cased unless -j, non-ACGT
// masked (nucleotide: rand() & 3; color: previous base), color = XOR of adjacent bases.
int load_reference(const char* path, const storm_params& p, std::vector<HostRef>& out, std::string& err) {
  std::string data;
  if (!read_file(path, data)) { err = std::string("cannot read ") + path; return STORM_ERR_IO; }
  out.clear();
  GlibcRand rnd(1);
  size_t i = 0, n = data.size();
  bool any_header = false;
  while (i < n) {
    HostRef ref;
    // header line (first line of a record)
    if (data[i] == '>') {
      any_header = true;
      size_t e = data.find('\n', i);
      if (e == std::string::npos) e = n;
      size_t nl = strcspn(data.c_str() + i + 1, " \r\n");
      if (nl == 0) { err = "empty fasta header line in the reference sequence"; return STORM_ERR_INVALID; }
      if (nl > e - i - 1) nl = e - i - 1;
      ref.name = data.substr(i + 1, nl);
      if (!ref.name.empty() && (ref.name[0] == '=' || ref.name[0] == '*')) ref.name[0] = '-';
      i = e + 1;
    } else if (any_header || !out.empty()) {
      break;
    }
    uint8_t code_prev = 0;
    int64_t mstart = -1;
    for (; i < n; ++i) {
      int c = (unsigned char)data[i];
      if (!p.ignore_lowercase) c = toupper(c);
      if (isspace(c)) continue;
      if (c == '>') break;
      uint8_t code;
      const int64_t pos = (int64_t)ref.code.size();
      bool masked = !(c == 'A' || c == 'C' || c == 'G' || c == 'T');
      if (masked) {
        code = p.color ? code_prev : (uint8_t)(rnd.next() & 3);
        if (mstart == -1) mstart = pos;
      } else {
        code = base_symbol_code(c);
        if (mstart != -1) { ref.regions.emplace_back((int)mstart, (int)pos); mstart = -1; }
      }
      if (!p.color) {
        ref.code.push_back(code);
        ref.mask.push_back(masked ? 1 : 0);
      } else {
        if (ref.first_base == 0xff) {
          // reference_data.c:121 marks the mask at the current index even for the first base
          if (masked) { if (ref.mask.empty()) ref.mask.push_back(1); else ref.mask[0] = 1; }
          code_prev = ref.first_base = code;
        } else {
          ref.code.push_back((uint8_t)(code_prev ^ code));
          if (ref.mask.size() < ref.code.size()) ref.mask.push_back(masked ? 1 : 0);
          else if (masked) ref.mask[ref.code.size() - 1] = 1;
          code_prev = code;
        }
      }
    }
    if (mstart != -1) ref.regions.emplace_back((int)mstart, (int)ref.code.size());
    ref.size = (int64_t)ref.code.size();
    ref.mask.resize(ref.code.size(), 0);
    if (ref.size == 0) {
      if (out.empty()) { err = "empty fasta chunk in the reference sequence"; return STORM_ERR_INVALID; }
      break;
    }
    out.push_back(std::move(ref));
    if (!any_header) break;
  }
  if (out.empty()) { err = "reference could not be loaded"; return STORM_ERR_INVALID; }
  return STORM_OK;
}

// ------------------------------------------------------------------------------------ reads
static inline int quality_level(int q) {   // read_quality.c:10-16
  static const short ub[4] = {2, 7, 10, 93};
  int i = 0;
  while (i < 3 && ub[i] < q) ++i;
  return i;
}

static inline void put2(uint8_t* s, int64_t k, int v) { s[k >> 2] |= (uint8_t)((v & 3) << ((k & 3) << 1)); }
static inline int get2(const uint8_t* s, int64_t k) { return (s[k >> 2] >> ((k & 3) << 1)) & 3; }

namespace {
struct RawRead {
  std::string name;
  std::string seq;     // sequence line(s), '\r\n' removed
  std::string qual;    // FASTQ quality line, or .qual integers line
};
}  // namespace

static std::string parse_read_name(const char* line, size_t len) {   // reads_data.c:460-474
  size_t k = 1;
  while (k < len && (unsigned char)line[k] > ' ' && (unsigned char)line[k] <= '~' && line[k] != '@') {
    ++k;
    if (k >= 256) break;
  }
  return std::string(line + 1, k - 1);
}

// reads_data.c:538-564: common printable prefix of the first and last names
static std::string common_prefix(const std::vector<std::string>& names) {
  if (names.empty()) return "";
  const std::string &a = names.front(), &b = names.back();
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

int load_reads(const char* path, const char* qual_path, const storm_params& p, HostReads& out, std::string& err) {
  std::string data;
  if (!read_file(path, data)) { err = std::string("cannot read ") + path; return STORM_ERR_IO; }
  std::vector<std::pair<size_t, size_t>> lines;
  split_lines(data, lines);
  std::vector<RawRead> raw;
  std::string title;
  const char* dot = strrchr(path, '.');
  const bool fastq = dot && (!strcmp(dot, ".fastq") || !strcmp(dot, ".fq") || !strcmp(dot, ".fastq2") || !strcmp(dot, ".fq2"));
  size_t li = 0;
  // leading comments, "# Title: " names the read group (reads_data.c:380-390)
  while (li < lines.size() && lines[li].second > 0 && data[lines[li].first] == '#') {
    if (title.empty() && lines[li].second >= 9 && data.compare(lines[li].first, 9, "# Title: ") == 0)
      title = data.substr(lines[li].first + 9, lines[li].second - 9);
    ++li;
  }
  if (fastq) {   // reads_data.c:351-566
    // get_input_line_type_fastq (load_data.c:85-107): lines are typed by their position relative to
    // the first line after the comments and by their first character; crt is the type of the
    // previous typed line (an empty line leaves it unchanged)
    enum { LT_UNK, LT_COM, LT_HEAD, LT_SEQ, LT_QHEAD, LT_QUAL };
    const size_t li0 = li;
    int crt = LT_UNK;
    auto type = [&](size_t j) {
      if (lines[j].second == 0) return (int)LT_UNK;
      const char c0 = data[lines[j].first];
      const size_t rel = (j - li0) % 4;
      int t;
      if (crt == LT_COM || c0 == '#') t = LT_COM;
      else if (rel == 0 && c0 == '@') t = LT_HEAD;
      else if (rel == 2 && c0 == '+') t = LT_QHEAD;
      else if (rel == 1 && crt == LT_HEAD) t = LT_SEQ;
      else if (rel == 3 && crt == LT_QHEAD) t = LT_QUAL;
      else t = LT_UNK;
      crt = (t == LT_SEQ || t == LT_HEAD || t == LT_QUAL || t == LT_QHEAD) ? t : LT_UNK;
      return t;
    };
    // first pass (reads_data.c:395-434): a header must be followed by a sequence line, and an
    // untyped line (empty, or out of place) makes the file corrupt
    int64_t count = 0;
    int next = -1;
    for (size_t j = li; j < lines.size();) {
      const int t = type(j++);
      if (t == LT_HEAD && j < lines.size()) {
        next = type(j++);
        if (next == LT_SEQ) { ++count; continue; }
      }
      if ((t == LT_HEAD && next != LT_SEQ) || t == LT_UNK) {
        err = "reads file seems corrupt at line " + std::to_string(j);
        return STORM_ERR_INVALID;
      }
    }
    // second pass (reads_data.c:452-532): a header starts a read, the sequence and quality lines
    // fill it (a quality line typed as a comment leaves the read without quality)
    raw.reserve((size_t)count);
    for (size_t j = li; j < lines.size() && (int64_t)raw.size() <= count; ++j) {
      const int t = type(j);
      if (t == LT_HEAD) {
        if ((int64_t)raw.size() == count) break;
        RawRead r;
        r.name = parse_read_name(data.c_str() + lines[j].first, lines[j].second);
        raw.push_back(std::move(r));
      } else if (t == LT_SEQ && !raw.empty()) {
        raw.back().seq = data.substr(lines[j].first, lines[j].second);
      } else if (t == LT_QUAL && !raw.empty()) {
        raw.back().qual = data.substr(lines[j].first, lines[j].second);
      }
    }
  } else {       // reads_data.c:73-349, '>' headers, sequence lines until the next header
    while (li < lines.size()) {
      if (lines[li].second == 0) { ++li; continue; }
      const char* h = data.c_str() + lines[li].first;
      if (h[0] != '>') { ++li; continue; }
      RawRead r;
      r.name = parse_read_name(h, lines[li].second);
      ++li;
      while (li < lines.size() && lines[li].second > 0 && data[lines[li].first] != '>' && data[lines[li].first] != '#') {
        r.seq.append(data, lines[li].first, lines[li].second);
        ++li;
      }
      raw.push_back(std::move(r));
    }
    if (qual_path) {   // .qual: one line of integers per read (reads_data.c:297-326)
      std::string qd;
      if (!read_file(qual_path, qd)) { err = std::string("cannot read ") + qual_path; return STORM_ERR_IO; }
      std::vector<std::pair<size_t, size_t>> ql;
      split_lines(qd, ql);
      size_t k = 0;
      for (auto& l : ql) {
        if (k >= raw.size()) break;
        if (l.second == 0) continue;
        char c = qd[l.first];
        if (c == '#' || c == '>') continue;
        raw[k++].qual = qd.substr(l.first, l.second);
      }
    }
  }
  const int64_t n = (int64_t)raw.size();
  if (n == 0) { err = "no read loaded"; return STORM_ERR_INVALID; }
  out = HostReads();
  out.n = n;
  out.names.resize(n);
  out.len.resize(n);
  int maxL = 0, minL = 1 << 30;
  for (int64_t i = 0; i < n; ++i) {
    int L = (int)raw[i].seq.size() - (p.color ? 2 : 0);
    if (L <= 0) { err = "read #" + std::to_string(i + 1) + " is empty"; return STORM_ERR_INVALID; }
    out.len[i] = L;
    maxL = std::max(maxL, L);
    minL = std::min(minL, L);
  }
  if (p.strict_length && minL != maxL) {
    err = "reads have different lengths (the reference accepts one length per run)";
    return STORM_ERR_INVALID;
  }
  if (maxL > 256) { err = "reads longer than 256 codes are not supported"; return STORM_ERR_INVALID; }
  out.max_len = maxL;
  out.code_stride = ((maxL + 3) / 4 + 7) & ~7;
  out.qual_stride = out.code_stride;
  out.codes.assign((size_t)(n * out.code_stride), 0);
  out.has_qual = false;
  for (auto& r : raw) if (!r.qual.empty()) { out.has_qual = true; break; }
  if (out.has_qual) out.qual.assign((size_t)(n * out.qual_stride), 0);
  out.tag.resize(n);
  if (p.color) { out.first_base.assign(n, 0); out.first_qual.assign(n, 0); }
  std::vector<short> q(maxL + 2);
  for (int64_t i = 0; i < n; ++i) {
    RawRead& r = raw[i];
    out.names[i] = r.name;
    const int L = out.len[i];
    uint8_t* cs = &out.codes[(size_t)(i * out.code_stride)];
    const char* s = r.seq.c_str();
    if (!p.color) {   // codes.c:28-44: uppercase ACGT only, anything else -> 0
      for (int k = 0; k < L; ++k) {
        char c = s[k];
        put2(cs, k, c == 'A' ? 0 : c == 'C' ? 1 : c == 'G' ? 2 : c == 'T' ? 3 : 0);
      }
    } else {          // reads_data.c:488-491: primer base + first color -> first base
      out.first_base[i] = (uint8_t)(base_symbol_code(s[0]) ^ (s[1] & 3));
      for (int k = 0; k < L; ++k) put2(cs, k, s[k + 2] & 3);
    }
    uint64_t tag = 0, ct = 0;   // reads_data.c:494-502
    for (int k = 0; k < L; ++k) {
      ct = (ct << 2) | (uint64_t)((((get2(cs, k) + 1) & 3)) ^ 1);
      if (ct > tag) tag = ct;
    }
    out.tag[i] = tag;
    if (!out.has_qual) continue;
    if (r.qual.empty()) {   // no quality: levels 3 everywhere, as the reference's NULL quality
      if (out.no_qual.empty()) out.no_qual.assign((size_t)n, 0);
      out.no_qual[(size_t)i] = 1;
      memset(&out.qual[(size_t)(i * out.qual_stride)], 0xFF, (size_t)out.qual_stride);
      if (p.color) out.first_qual[(size_t)i] = 3;
      continue;
    }
    uint8_t* qs = &out.qual[(size_t)(i * out.qual_stride)];
    std::fill(q.begin(), q.end(), 0);
    if (fastq) {   // reads_data.c:516-531
      const int lim = std::min(L, (int)r.qual.size());
      for (int k = p.color ? 1 : 0; k < lim; ++k) {
        int v = (unsigned char)r.qual[k] - p.qual_base;
        if (v < 0) { err = "quality character below the -b base"; return STORM_ERR_INVALID; }
        q[k] = (short)std::min(v, 93);
      }
      if (p.color) {
        int v0 = (unsigned char)r.qual[0] - p.qual_base;
        if (v0 < 0) { err = "quality character below the -b base"; return STORM_ERR_INVALID; }
        out.first_qual[i] = (uint8_t)quality_level(std::min(v0, 93));
        // compress_quality_sequence(shift 1) reads q[1..L]; q[L] lies past the reference's buffer
        // (reads_data.c:518-529) and is taken as 0 here
        for (int k = 0; k < L; ++k) put2(qs, k, quality_level(q[k + 1]));
      } else {
        for (int k = 0; k < L; ++k) put2(qs, k, quality_level(q[k]));
      }
    } else {       // integers separated by spaces, L+1 values in color mode (first = primer color)
      const int want = L + (p.color ? 1 : 0);
      const char* c = r.qual.c_str();
      for (int k = 0; k < want; ++k) {
        while (*c == ' ') ++c;
        if (!*c) break;
        q[k] = (short)atoi(c);
        while (*c && *c != ' ') ++c;
      }
      if (p.color) {
        out.first_qual[i] = (uint8_t)quality_level(q[0]);
        for (int k = 0; k < L; ++k) put2(qs, k, quality_level(q[k + 1]));
      } else {
        for (int k = 0; k < L; ++k) put2(qs, k, quality_level(q[k]));
      }
    }
  }
  out.db_name = title.empty() ? common_prefix(out.names) : title;
  return STORM_OK;
}

// reads_data.c:59-67 (glibc 2.35 qsort of this array size is a stable merge sort)
void sort_reads_by_tag(HostReads& r) {
  std::vector<int64_t> perm((size_t)r.n);
  std::iota(perm.begin(), perm.end(), 0);
  std::stable_sort(perm.begin(), perm.end(), [&](int64_t a, int64_t b) { return r.tag[a] < r.tag[b]; });
  HostReads s;
  s.db_name = r.db_name; s.n = r.n; s.max_len = r.max_len; s.has_qual = r.has_qual;
  s.code_stride = r.code_stride; s.qual_stride = r.qual_stride;
  s.names.resize(r.n); s.len.resize(r.n); s.tag.resize(r.n);
  s.codes.resize(r.codes.size());
  if (r.has_qual) s.qual.resize(r.qual.size());
  if (!r.first_base.empty()) { s.first_base.resize(r.n); s.first_qual.resize(r.n); }
  for (int64_t i = 0; i < r.n; ++i) {
    const int64_t j = perm[i];
    if (!r.no_qual.empty()) {
      if (s.no_qual.empty()) s.no_qual.assign((size_t)r.n, 0);
      s.no_qual[i] = r.no_qual[j];
    }
    s.names[i] = std::move(r.names[j]);
    s.len[i] = r.len[j];
    s.tag[i] = r.tag[j];
    memcpy(&s.codes[i * s.code_stride], &r.codes[j * r.code_stride], (size_t)s.code_stride);
    if (r.has_qual) memcpy(&s.qual[i * s.qual_stride], &r.qual[j * r.qual_stride], (size_t)s.qual_stride);
    if (!r.first_base.empty()) { s.first_base[i] = r.first_base[j]; s.first_qual[i] = r.first_qual[j]; }
  }
  r = std::move(s);
}

// ------------------------------------------------------------------------------------ index
// index.c:117-156: average and standard deviation of the key counts in double, summed in key
// order with the reference's expression order, then every key whose count exceeds
// avg + l * sd is marked in erased[].  `count` is the number of indexed windows.
int64_t repeat_filter(const int32_t* counts, int64_t ksize, int64_t count, int stdev_distance, uint8_t* erased) {
  int64_t n = 0;
  double avg = 0;
  for (int64_t k = 0; k < ksize; ++k)
    if (counts[k] > 0) avg += 1.0 / (int)count * counts[k] * counts[k];
  avg /= (int)count;
  double sd = 0;
  for (int64_t k = 0; k < ksize; ++k)
    if (counts[k] > 0) sd += (counts[k] - avg) * (counts[k] - avg) / (int)count;
  sd = sqrt(sd);
  for (int64_t k = 0; k < ksize; ++k)
    if (counts[k] && (counts[k] - avg) > stdev_distance * sd) { erased[k] = 1; n++; }
  return n;
}

// index.c:19-195 with the contiguous layout of index.c:157-188: positions q in [0, size-span]
// in ascending order per key, masked windows skipped (index.c:66-72), keys whose count exceeds
// avg + l*sd erased (index.c:117-156, same double expression order).
void build_index(const HostRef& ref, const SeedSpec& seed, const storm_params& p, HostIndex& out) {
  const int64_t ksize = 1LL << (2 * seed.weight);
  out.first.assign((size_t)ksize + 1, 0);
  out.map.clear();
  out.erased = 0;
  const int64_t map_size = ref.size - seed.span + 1;
  if (map_size <= 0 || seed.span >= ref.size) return;
  std::vector<int32_t> counts((size_t)ksize, 0);
  std::vector<int32_t> keys;
  std::vector<int32_t> posv;
  keys.reserve((size_t)map_size);
  posv.reserve((size_t)map_size);
  size_t r = 0;
  const uint8_t* code = ref.code.data();
  for (int64_t i = 0; i < map_size; ++i) {
    while (r < ref.regions.size() && i + seed.span >= ref.regions[r].first) {
      i = ref.regions[r].second - 1;
      ++r;
    }
    // a region reaching the end of the sequence puts i at or past map_size; the reference then
    // indexes a window that reads past its map[] (undefined), such positions are left out
    if (i >= map_size) break;
    int key = 0;
    for (int k = 0; k < seed.weight; ++k) {
      const int64_t q = i + seed.offs[k];
      key = (key << 2) | (q < ref.size ? code[q] : 0);
    }
    counts[key]++;
    keys.push_back(key);
    posv.push_back((int32_t)i);
  }
  const int64_t count = (int64_t)keys.size();
  std::vector<uint8_t> erased((size_t)ksize, 0);
  if (p.stdev_distance > 0) out.erased = repeat_filter(counts.data(), ksize, count, p.stdev_distance, erased.data());
  for (int64_t k = 0; k < ksize; ++k) out.first[k + 1] = out.first[k] + (erased[k] ? 0 : counts[k]);
  out.map.resize((size_t)out.first[ksize] + 1);
  std::vector<int32_t> fill(out.first.begin(), out.first.end() - 1);
  for (int64_t e = 0; e < count; ++e) {
    const int k = keys[e];
    if (!erased[k]) out.map[fill[k]++] = posv[e];
  }
}

// ------------------------------------------------------------------------------------ SAM
// alignment__traceback_to_CIGAR (alignment.c:749-887); tb holds the codes in alignment order
// (end to start), so the walk below reads them from the last one.
static int traceback_to_cigar(const HostSlot& s, const HostRef& ref, const std::vector<uint8_t>* sym, bool color,
                              std::string& cigar, std::string& seq) {
  static const char BASES[4] = {'A', 'C', 'G', 'T'};
  int ed = 0, span = color ? 1 : 0;
  int prev = 0x3;
  int64_t ref_idx = s.ref_start;
  uint8_t ref_hist = 0, read_hist = 0;
  std::vector<uint8_t> codes;
  codes.reserve((size_t)s.tb_len + 2);
  if (color) codes.push_back(0);
  cigar.clear();
  char buf[32];
  auto refcode = [&](int64_t k) -> int { return (k >= 0 && k < ref.size) ? ref.code[(size_t)k] : 0; };
  for (int i = 0; i < s.tb_len; ++i) {
    const int src = s.tb_len - 1 - i;
    const int cell = (s.tb[(size_t)src >> 1] >> ((src & 1) ? 0 : 4)) & 0xF;
    int type;
    if ((cell & 0xC) == 0x4) type = 0x4;
    else if (color && (cell & 0xC) == 0x8) type = 0x8;
    else if ((cell & 0xC) == 0xC) type = 0xC;
    else type = cell;
    uint8_t code = (uint8_t)(cell & 3);
    switch (type) {
      case 0x3:
        ed--;
        if (!color) code = (uint8_t)refcode(ref_idx);
        /* fall through */
      case 0x4:
        ed++;
        if (color) { ref_hist ^= (uint8_t)refcode(ref_idx); code = (uint8_t)(ref_hist ^ read_hist); read_hist ^= code; }
        codes.push_back(code);
        ref_idx++;
        break;
      case 0x8:
        ed++;
        codes.push_back(code);
        read_hist ^= code;
        ref_hist ^= (uint8_t)refcode(ref_idx);
        ref_idx++;
        break;
      case 0xC:
        ed++;
        codes.push_back(code);
        if (color) read_hist ^= code;
        break;
      case 0x1:
        ed++;
        if (color) ref_hist ^= (uint8_t)refcode(ref_idx);
        ref_idx++;
        break;
      default:
        break;
    }
    const int crt = (type == 0x1 || type == 0xC) ? type : 0x3;
    if (crt == prev) ++span;
    else {
      if (span) { snprintf(buf, sizeof buf, "%d%c", span, prev == 0x3 ? 'M' : prev == 0xC ? 'I' : prev == 0x1 ? 'D' : '?'); cigar += buf; }
      span = 1;
      prev = crt;
    }
  }
  if (span) { snprintf(buf, sizeof buf, "%d%c", span, prev == 0x3 ? 'M' : prev == 0xC ? 'I' : prev == 0x1 ? 'D' : '?'); cigar += buf; }
  seq.resize(codes.size());
  if (!color) {
    for (size_t u = 0; u < codes.size(); ++u) seq[u] = BASES[codes[u] & 3];
  } else {
    uint8_t c = (*sym)[(size_t)s.ref_start];
    for (size_t u = 0; u < codes.size(); ++u) { c = (uint8_t)(c ^ codes[u]); seq[u] = BASES[c & 3]; }
  }
  return ed;
}

SamFormat::SamFormat(const std::vector<HostRef>& refs_, const HostReads& reads_, const storm_params& p_)
    : refs(refs_), reads(reads_), p(p_), color(p_.color != 0) {
  if (color) {   // hit_map.c:236-249, base-space symbols of each color reference (zeroed buffer)
    sym.resize(refs.size());
    for (size_t r = 0; r < refs.size(); ++r) {
      sym[r].resize((size_t)refs[r].size + 1);
      uint8_t s = refs[r].first_base;
      for (int64_t k = 0; k < refs[r].size; ++k) { sym[r][(size_t)k] = s; s = (uint8_t)(s ^ refs[r].code[(size_t)k]); }
      sym[r][(size_t)refs[r].size] = s;
    }
  }
}

void sam_header(FILE* out, const std::vector<HostRef>& refs, bool color, const std::string& db_name,
                const char* sort_order, int argc, char** argv) {   // hit_map.c:297-315
  fprintf(out, "@HD\tVN:1.4\tSO:%s\n", sort_order);
  for (auto& r : refs)
    fprintf(out, "@SQ\tSN:%s\tLN:%ld\n", r.name.empty() ? "*" : r.name.c_str(), (long)(r.size + (color ? 1 : 0)));
  const char* rg = db_name.empty() ? "*" : db_name.c_str();
  fprintf(out, "@RG\tID:%s\tSM:%s\n", rg, rg);
  fprintf(out, "@PG\tID:SToRM\tVN:0.0101\tCL:\"");
  for (int a = 0; a < argc; ++a) fprintf(out, "%s ", argv[a]);
  fprintf(out, "\"\n");
}

void SamFormat::header(FILE* out, const char* sort_order, int argc, char** argv) const {
  sam_header(out, refs, color, reads.db_name, sort_order, argc, argv);
}

int sam_mapq(int diff, int delta_m) {   // hit_map.c:374-399, n = diff / (m - x)
  const double lg = log(.3);
  const double nn = (double)diff / (double)delta_m;
  int mq = -4.343 * (nn * lg);
  mq = (int)(mq - 4.343 * lg);
  mq = mq >= 0 ? mq : 0;
  return mq < 254 ? mq : 254;
}

int SamFormat::mapq(const HostSlot* row, int k) const {   // rank 0 against max(rank 1, t), later ranks against rank 0
  const int s1 = row[0].score, s2 = (k == 0) ? std::max(row[1].score, p.t) : row[k].score;
  return sam_mapq(k == 0 ? s1 - s2 : s2 - s1, p.match - p.mismatch);
}

void SamFormat::record(std::string& line, int64_t id, const HostSlot* row, int k, int flag, std::string& cigar,
                       std::string& seq) const {   // hit_map.c:401-448
  static const short UB[4] = {2, 7, 10, 93};
  char buf[96];
  const int L = reads.len[id];
  const bool rq = reads.has_qual && !(!reads.no_qual.empty() && reads.no_qual[(size_t)id]);
  const uint8_t* qs = rq ? &reads.qual[(size_t)(id * reads.qual_stride)] : nullptr;
  const HostRef& ref = refs[(size_t)row[k].ref_id];
  const int ed = traceback_to_cigar(row[k], ref, color ? &sym[(size_t)row[k].ref_id] : nullptr, color, cigar, seq);
  line += reads.names[(size_t)id];
  snprintf(buf, sizeof buf, "\t%d\t", flag);
  line += buf;
  line += ref.name.empty() ? "*" : ref.name;
  snprintf(buf, sizeof buf, "\t%d\t%d\t", row[k].ref_start + 1, mapq(row, k));
  line += buf;
  line += cigar;
  line += "\t*\t0\t0\t";
  line += seq;
  line += '\t';
  if (qs) {
    auto qc = [&](int lv) { int v = p.qual_base + UB[lv]; return (char)(v < 126 ? v : 126); };
    if (row[k].sense == -1) {
      for (int q = L - 1; q >= 0; --q) line += qc(get2(qs, q));
      if (color) line += qc(reads.first_qual[(size_t)id]);
    } else {
      if (color) line += qc(reads.first_qual[(size_t)id]);
      for (int q = 0; q < L; ++q) line += qc(get2(qs, q));
    }
  } else {
    line += '*';
  }
  snprintf(buf, sizeof buf, "\tAS:i:%d\tNM:i:%d", row[k].score, ed);
  line += buf;
}

bool write_blocks(FILE* out, int64_t n, const std::function<void(int64_t, std::string&, std::string&, std::string&)>& fmt) {
  bool ok = true;
  const int nt = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
  const int64_t per_thread = 32768, block = per_thread * nt;
  std::vector<std::string> part((size_t)nt);
  for (int64_t b0 = 0; b0 < n; b0 += block) {
    const int64_t b1 = std::min(n, b0 + block);
    auto work = [&](int t) {
      std::string& out_t = part[(size_t)t];
      out_t.clear();
      std::string cigar, seq;
      const int64_t r0 = b0 + t * per_thread, r1 = std::min(b1, r0 + per_thread);
      for (int64_t id = r0; id < r1; ++id) fmt(id, out_t, cigar, seq);
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nt && b0 + t * per_thread < b1; ++t) th.emplace_back(work, t);
    work(0);
    for (auto& x : th) x.join();
    for (int t = 0; t < nt; ++t)
      if (!part[(size_t)t].empty() && fwrite(part[(size_t)t].data(), 1, part[(size_t)t].size(), out) != part[(size_t)t].size())
        ok = false;
  }
  return ok && !ferror(out);
}

int write_unmapped_fastq(const char* path, const HostReads& reads, const std::vector<HostSlot>& slots,
                         const storm_params& p, std::string& err) {
  FILE* out = fopen(path, "wb");
  if (!out) { err = std::string("cannot write ") + path; return STORM_ERR_IO; }
  static const char BASES[4] = {'A', 'C', 'G', 'T'};
  static const short UB[4] = {2, 7, 10, 93};
  const bool color = p.color != 0;
  const char qmax = (char)std::min(p.qual_base + 93, 126);   // MAX_READ_QUALITY (read_quality.h:8)
  const bool ok = write_blocks(out, reads.n, [&](int64_t id, std::string& line, std::string&, std::string&) {
    if (slots[(size_t)(4 * id)].ref_id != -1) return;
    const int L = reads.len[id];
    const std::string& nm = reads.names[(size_t)id];
    line += '@';
    line += nm;
    line += '\n';
    const uint8_t* cs = &reads.codes[(size_t)(id * reads.code_stride)];
    if (color) {
      line += BASES[reads.first_base[(size_t)id] & 3];
      for (int k = 0; k < L; ++k) line += (char)('0' + get2(cs, k));
    } else {
      for (int k = 0; k < L; ++k) line += BASES[get2(cs, k)];
    }
    line += "\n+";
    line += nm;
    line += '\n';
    const bool rq = reads.has_qual && !(!reads.no_qual.empty() && reads.no_qual[(size_t)id]);
    if (rq) {
      const uint8_t* qs = &reads.qual[(size_t)(id * reads.qual_stride)];
      auto qc = [&](int lv) { int v = p.qual_base + UB[lv]; return (char)(v < 126 ? v : 126); };
      if (color) line += qc(reads.first_qual[(size_t)id]);
      for (int q = 0; q < L; ++q) line += qc(get2(qs, q));
    } else {
      line.append((size_t)(L + (color ? 1 : 0)), qmax);
    }
    line += '\n';
  });
  if (fclose(out) != 0 || !ok) { err = std::string("error writing ") + path; return STORM_ERR_IO; }
  return STORM_OK;
}

}  // namespace storm
