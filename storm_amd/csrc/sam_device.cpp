// sam_device.cpp -- the -M SAM records of a map call formatted on the device (k_sam.hip) from the
// hit map the call left there: the host uploads the read names and the tables (once per genome:
// planes of every sequence, names, color base symbols; per score setting: the MAPQ table) and
// writes the text the device returns, in the caller's read order, in chunks of reads.
#include <algorithm>
#include <cstring>

#include "mapper.h"

namespace storm {

namespace {

constexpr int64_t SAM_CHUNK = 1 << 20;   // reads formatted per launch pair

int sam_tables(storm_mapper* m) {
  SamTables& T = m->res->sam;
  std::lock_guard<std::mutex> lk(T.mu);
  if (T.ready) return STORM_OK;
  int rc = ensure_planes(m);
  if (rc) return rc;
  const auto& refs = m->res->refs;
  const size_t nr = refs.size();
  std::vector<GenomeDev> gd(nr);
  std::vector<int32_t> noff(nr + 1, 0);
  std::string names;
  std::vector<int64_t> soff(nr + 1, 0);
  std::vector<uint8_t> ck;
  for (size_t r = 0; r < nr; ++r) {
    gd[r] = m->res->rstate[r]->g.view();
    names += refs[r].name.empty() ? std::string("*") : refs[r].name;
    noff[r + 1] = (int32_t)names.size();
    if (m->p.color) {   // base symbol before every 32-code word (hit_map.c:236-249)
      const HostRef& h = refs[r];
      uint8_t s = h.first_base;
      for (int64_t k = 0; k < h.size; ++k) {
        if ((k & 31) == 0) ck.push_back((uint8_t)(s & 3));
        s = (uint8_t)(s ^ h.code[(size_t)k]);
      }
      ck.push_back((uint8_t)(s & 3));   // a word past the end (a hit never starts there)
      soff[r + 1] = (int64_t)ck.size();
    }
  }
  HIPOK(T.genomes.ensure(nr * sizeof(GenomeDev)));
  HIPOK(T.name_off.ensure((nr + 1) * 4));
  HIPOK(T.names.ensure(names.size() + 1));
  HIPOK(hipMemcpyAsync(T.genomes.p, gd.data(), nr * sizeof(GenomeDev), hipMemcpyHostToDevice, m->st));
  HIPOK(hipMemcpyAsync(T.name_off.p, noff.data(), (nr + 1) * 4, hipMemcpyHostToDevice, m->st));
  HIPOK(hipMemcpyAsync(T.names.p, names.data(), names.size(), hipMemcpyHostToDevice, m->st));
  if (m->p.color) {
    HIPOK(T.sym_ck.ensure(ck.size()));
    HIPOK(T.sym_off.ensure((nr + 1) * 8));
    HIPOK(hipMemcpyAsync(T.sym_ck.p, ck.data(), ck.size(), hipMemcpyHostToDevice, m->st));
    HIPOK(hipMemcpyAsync(T.sym_off.p, soff.data(), (nr + 1) * 8, hipMemcpyHostToDevice, m->st));
  }
  HIPOK(hipStreamSynchronize(m->st));   // the host vectors are released on return
  T.ready = true;
  return STORM_OK;
}

}  // namespace

SamRows sam_rows_of_view(const storm_reads_view* v, std::string& names, std::vector<int64_t>& name_off) {
  SamRows r{};
  r.n = v->n_reads;
  if (v->names) {
    name_off.resize((size_t)r.n + 1);
    size_t total = 0;
    for (int64_t i = 0; i < r.n; ++i) total += strlen(v->names[i]);
    names.resize(total);
    size_t o = 0;
    for (int64_t i = 0; i < r.n; ++i) {
      name_off[(size_t)i] = (int64_t)o;
      const size_t l = strlen(v->names[i]);
      memcpy(&names[o], v->names[i], l);
      o += l;
    }
    name_off[(size_t)r.n] = (int64_t)o;
    r.names = names.data();
    r.name_off = name_off.data();
  }
  r.no_qual = v->qual ? v->no_qual : nullptr;
  r.has_qual = v->qual != nullptr;
  r.has_first_qual = v->first_qual != nullptr;
  r.code_stride = v->code_stride;
  r.qual_stride = v->qual ? v->qual_stride : 0;
  return r;
}

int sam_device(storm_mapper* m, const SamRows& rows, const FinishCtx& fc, const SamSink& sink, bool with_offsets,
               int64_t& mapped) {
  const int64_t n = fc.n;
  mapped = 0;
  if (n == 0) return STORM_OK;
  if (rows.n != n) return fail(STORM_ERR_INVALID, "SAM rows (%lld) differ from the reads mapped (%lld)", (long long)rows.n,
                               (long long)n);
  int rc = sam_tables(m);
  if (rc) return rc;
  hipStream_t st = m->st;
  // MAPQ by score difference (the reference's double expression, libm log on the host)
  const int dm = m->p.match - m->p.mismatch;
  if (m->sam_lut_dm != dm || !m->sam_lut.p) {
    std::vector<uint8_t> lut(65536);
    for (int d = 0; d < 65536; ++d) lut[(size_t)d] = (uint8_t)sam_mapq(d - 32768, dm);
    HIPOK(m->sam_lut.ensure(lut.size()));
    HIPOK(hipMemcpyAsync(m->sam_lut.p, lut.data(), lut.size(), hipMemcpyHostToDevice, st));
    HIPOK(hipStreamSynchronize(st));
    m->sam_lut_dm = dm;
  }
  // caller row -> device row
  const std::vector<int64_t>& order = *fc.order;
  bool identity = true;
  for (int64_t i = 0; i < n && identity; ++i) identity = order[(size_t)i] == i;
  std::vector<int32_t> inv;
  if (!identity) {
    inv.resize((size_t)n);
    for (int64_t i = 0; i < n; ++i) inv[(size_t)order[(size_t)i]] = (int32_t)i;
    HIPOK(m->sam_inv.ensure((size_t)n * 4));
    HIPOK(hipMemcpyAsync(m->sam_inv.p, inv.data(), (size_t)n * 4, hipMemcpyHostToDevice, st));
  }
  if (rows.names) {
    const int64_t nb = rows.name_off[n];
    HIPOK(m->sam_names.ensure((size_t)nb + 1));
    HIPOK(m->sam_name_off.ensure((size_t)(n + 1) * 8));
    if (nb) HIPOK(hipMemcpyAsync(m->sam_names.p, rows.names, (size_t)nb, hipMemcpyHostToDevice, st));
    HIPOK(hipMemcpyAsync(m->sam_name_off.p, rows.name_off, (size_t)(n + 1) * 8, hipMemcpyHostToDevice, st));
  }
  if (rows.no_qual) {
    HIPOK(m->sam_no_qual.ensure((size_t)n));
    HIPOK(hipMemcpyAsync(m->sam_no_qual.p, rows.no_qual, (size_t)n, hipMemcpyHostToDevice, st));
  }
  HIPOK(m->sam_mapped.ensure(8));
  HIPOK(hipMemsetAsync(m->sam_mapped.p, 0, 8, st));
  const SamTables& T = m->res->sam;
  const DevReads& d = *fc.reads;
  SamArgs a{};
  a.slots = fc.slots;
  a.cells = fc.cells;
  a.inv = identity ? nullptr : m->sam_inv.as<int32_t>();
  a.codes = d.codes.as<uint8_t>();
  a.code_stride = rows.code_stride;
  a.qual = rows.has_qual ? d.qual.as<uint8_t>() : nullptr;
  a.qual_stride = rows.has_qual ? rows.qual_stride : 0;
  a.first_qual = (m->p.color && rows.has_first_qual) ? d.first_qual.as<uint8_t>() : nullptr;
  a.len = d.len.as<int32_t>();
  a.no_qual = rows.no_qual ? m->sam_no_qual.as<uint8_t>() : nullptr;
  a.names = rows.names ? m->sam_names.as<char>() : nullptr;
  a.name_off = rows.names ? m->sam_name_off.as<int64_t>() : nullptr;
  a.name_base = 0;
  a.genomes = T.genomes.as<GenomeDev>();
  a.sym_ck = m->p.color ? T.sym_ck.as<uint8_t>() : nullptr;
  a.sym_off = m->p.color ? T.sym_off.as<int64_t>() : nullptr;
  a.ref_names = T.names.as<char>();
  a.ref_name_off = T.name_off.as<int32_t>();
  a.mapq_lut = m->sam_lut.as<uint8_t>();
  a.M = m->p.max_hits;
  a.uniqueness = m->p.uniqueness;
  a.print_all = m->p.print_all;
  a.qual_base = m->p.qual_base;
  a.color = m->p.color ? 1 : 0;
  a.t = m->p.t;
  a.mapped = m->sam_mapped.as<int64_t>();
  const int64_t ch = std::min<int64_t>(n, SAM_CHUNK);
  HIPOK(m->sam_lens.ensure((size_t)(ch + 1) * 8));
  HIPOK(m->sam_offs.ensure((size_t)(ch + 1) * 8));
  size_t tmpb = 0;
  HIPOK(scan_exclusive_i64(nullptr, nullptr, ch + 1, nullptr, tmpb, st));
  HIPOK(m->ps.cub_tmp.ensure(tmpb + 256));
  std::vector<int64_t> offs;
  for (int64_t r0 = 0; r0 < n; r0 += ch) {
    const int64_t c = std::min(ch, n - r0);
    a.row0 = r0;
    a.n = c;
    a.lens = m->sam_lens.as<int64_t>();
    HIPOK(hipMemsetAsync(a.lens + c, 0, 8, st));
    HIPOK(launch_sam(a, false, st));
    size_t tb = m->ps.cub_tmp.n;
    HIPOK(scan_exclusive_i64(a.lens, m->sam_offs.as<int64_t>(), c + 1, m->ps.cub_tmp.p, tb, st));
    if (with_offsets) {
      offs.resize((size_t)c + 1);
      HIPOK(hipMemcpyAsync(offs.data(), m->sam_offs.p, (size_t)(c + 1) * 8, hipMemcpyDeviceToHost, st));
    }
    int64_t total = 0;
    HIPOK(hipMemcpyAsync(&total, m->sam_offs.as<int64_t>() + c, 8, hipMemcpyDeviceToHost, st));
    HIPOK(hipStreamSynchronize(st));
    if (total > 0) {
      HIPOK(m->sam_text.ensure((size_t)total));
      HIPOK(m->sam_host.ensure((size_t)total));
      a.offs = m->sam_offs.as<int64_t>();
      a.text = m->sam_text.as<char>();
      HIPOK(launch_sam(a, true, st));
      HIPOK(hipMemcpyAsync(m->sam_host.p, m->sam_text.p, (size_t)total, hipMemcpyDeviceToHost, st));
      HIPOK(hipStreamSynchronize(st));
    }
    rc = sink((const char*)m->sam_host.p, total, r0, c, with_offsets ? offs.data() : nullptr);
    if (rc) return rc;
  }
  HIPOK(hipMemcpyAsync(&mapped, m->sam_mapped.p, 8, hipMemcpyDeviceToHost, st));
  HIPOK(hipStreamSynchronize(st));
  return STORM_OK;
}

}  // namespace storm
