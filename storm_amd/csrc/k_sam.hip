// k_sam.hip -- the -M SAM records of a mapping run, formatted on the device.
//
// Semantics: the per-read part of hit_map__generate_SAM_output (hit_map.c:325-504) with
// alignment__traceback_to_CIGAR (alignment.c:749-887): for each read in the caller's order, its
// reportable ranks (k < -M, a hit, and the -u uniqueness margin, hit_map.c:359-370) as
//   QNAME FLAG RNAME POS MAPQ CIGAR * 0 0 SEQ QUAL AS:i NM:i NH:i
// (FLAG 16 on the reverse strand, 256 past rank 0; POS = ref_start + 1; MAPQ from a table the host
// builds with the reference's double expression, hit_map.c:374-399; SEQ rebuilt from the
// traceback and the genome -- in color space translated to bases from the reference's base
// symbols; QUAL the quantized quality-level bounds, reversed on the reverse strand), and with -A
// the unmapped reads' lines.  One thread formats one read.  Two launches share one body: the first
// counts each read's bytes (an exclusive scan then places the reads), the second writes them.
#include "storm_kernels.h"

namespace storm {

namespace {

// byte sink: counts (W = false) or stores (W = true) at dst
template <bool W>
struct Sink {
  char* dst;
  int64_t n;
  __device__ void put(char c) {
    if (W) dst[n] = c;
    ++n;
  }
  __device__ void str(const char* s, int len) {
    for (int i = 0; i < len; ++i) put(s[i]);
  }
  template <int N>
  __device__ void lit(const char (&s)[N]) {   // a string literal without its terminating NUL
    str(s, N - 1);
  }
  __device__ void num(int v) {   // decimal, as %d
    if (v < 0) { put('-'); v = -v; }
    char buf[12];
    int k = 0;
    do { buf[k++] = (char)('0' + v % 10); v /= 10; } while (v);
    while (k) put(buf[--k]);
  }
  __device__ void num64(int64_t v) {
    if (v < 0) { put('-'); v = -v; }
    char buf[21];
    int k = 0;
    do { buf[k++] = (char)('0' + v % 10); v /= 10; } while (v);
    while (k) put(buf[--k]);
  }
};

__device__ __forceinline__ int get2(const uint8_t* p, int i) { return (p[i >> 2] >> ((i & 3) << 1)) & 3; }

// genome code at position k of a sequence, 0 outside it (the reference buffers' zero padding)
__device__ __forceinline__ int refcode(const GenomeDev& g, int64_t k) {
  if (k < 0 || k >= g.size) return 0;
  return gcode(g, k);
}

// the base symbol before position k of a color sequence: the checkpoint of k's 32-code word xor
// the colors from the word start to k (parity of each plane's bits)
__device__ __forceinline__ int sym_at(const GenomeDev& g, const uint8_t* ck, int64_t k) {
  const int64_t w = k >> 5;
  const int r = (int)(k & 31);
  int s = ck[w];
  if (r) {
    const uint2 p = g.lh[(k + GPAD) >> 5];   // GPAD is a multiple of 32: the same word boundaries
    const uint32_t m = (1u << r) - 1u;
    s ^= (__popc(p.x & m) & 1) | ((__popc(p.y & m) & 1) << 1);
  }
  return s;
}

// alignment__traceback_to_CIGAR (alignment.c:749-887) over the codes of one traceback (stored end
// first): pass 0 emits the CIGAR, pass 1 the SEQ; returns the edit count (NM)
template <bool W>
__device__ int walk_tb(Sink<W>& out, const uint8_t* tb, int tb_len, const GenomeDev& g, const uint8_t* ck,
                       int64_t ref_start, bool color, bool seq) {
  static const char BASES[4] = {'A', 'C', 'G', 'T'};
  int ed = 0, span = color ? 1 : 0, prev = 0x3;
  int64_t ref_idx = ref_start;
  int ref_hist = 0, read_hist = 0;
  int c = 0;
  if (seq && color) {   // the leading symbol (the color SEQ starts from the reference's base)
    c = sym_at(g, ck, ref_start);
    out.put(BASES[c & 3]);
  }
  for (int i = 0; i < tb_len; ++i) {
    const int src = tb_len - 1 - i;
    const int cell = (tb[src >> 1] >> ((src & 1) ? 0 : 4)) & 0xF;
    int type;
    if ((cell & 0xC) == 0x4) type = 0x4;
    else if (color && (cell & 0xC) == 0x8) type = 0x8;
    else if ((cell & 0xC) == 0xC) type = 0xC;
    else type = cell;
    int code = cell & 3;
    bool emit = false;
    switch (type) {
      case 0x3:
        ed--;
        if (!color) code = refcode(g, ref_idx);
        // fall through
      case 0x4:
        ed++;
        if (color) { ref_hist ^= refcode(g, ref_idx); code = ref_hist ^ read_hist; read_hist ^= code; }
        emit = true;
        ref_idx++;
        break;
      case 0x8:
        ed++;
        emit = true;
        read_hist ^= code;
        ref_hist ^= refcode(g, ref_idx);
        ref_idx++;
        break;
      case 0xC:
        ed++;
        emit = true;
        if (color) read_hist ^= code;
        break;
      case 0x1:
        ed++;
        if (color) ref_hist ^= refcode(g, ref_idx);
        ref_idx++;
        break;
      default:
        break;
    }
    if (seq) {
      if (emit) {
        if (color) { c ^= code; out.put(BASES[c & 3]); }
        else out.put(BASES[code & 3]);
      }
    } else {
      const int crt = (type == 0x1 || type == 0xC) ? type : 0x3;
      if (crt == prev) ++span;
      else {
        if (span) { out.num(span); out.put(prev == 0x3 ? 'M' : prev == 0xC ? 'I' : prev == 0x1 ? 'D' : '?'); }
        span = 1;
        prev = crt;
      }
    }
  }
  if (!seq && span) { out.num(span); out.put(prev == 0x3 ? 'M' : prev == 0xC ? 'I' : prev == 0x1 ? 'D' : '?'); }
  return ed;
}

template <bool W>
__device__ void format_read(const SamArgs& a, int64_t v, Sink<W>& out) {
  static const char BASES[4] = {'A', 'C', 'G', 'T'};
  static const short UB[4] = {2, 7, 10, 93};   // read_quality.h:12-15
  const int64_t i = a.inv ? (int64_t)a.inv[v] : v;   // device row of caller row v
  const Slot* row = a.slots + i * 4;
  const int L = a.len[i];
  const bool rq = a.qual && !(a.no_qual && a.no_qual[v]);
  const uint8_t* qs = rq ? a.qual + i * a.qual_stride : nullptr;
  const int M = a.M;
  auto name = [&]() {
    if (a.names) out.str(a.names + a.name_off[v], (int)(a.name_off[v + 1] - a.name_off[v]));
    else { out.put('r'); out.num64(a.name_base + v); }
  };
  auto qc = [&](int lv) { const int x = a.qual_base + UB[lv]; out.put((char)(x < 126 ? x : 126)); };
  auto reportable = [&](int k) {
    return k < M && row[k].ref_start != -1 &&
           (a.uniqueness == 0 || ((k + 1 < M) && (row[k].score - row[k + 1].score) >= a.uniqueness) || (k + 1 == M));
  };
  if (row[0].ref_id != -1) {
    int nrep = 0;
    while (nrep < 4 && reportable(nrep)) ++nrep;
    for (int k = 0; k < nrep; ++k) {
      const Slot& s = row[k];
      const GenomeDev g = a.genomes[s.ref_id];
      const uint8_t* ck = a.color ? a.sym_ck + a.sym_off[s.ref_id] : nullptr;
      const uint8_t* tb = a.cells + (i * 4 + s.cell) * (int64_t)TBB;
      name();
      out.put('\t');
      out.num((s.sense == -1 ? 0x10 : 0) | (k ? 0x100 : 0));
      out.put('\t');
      out.str(a.ref_names + a.ref_name_off[s.ref_id], a.ref_name_off[s.ref_id + 1] - a.ref_name_off[s.ref_id]);
      out.put('\t');
      out.num(s.ref_start + 1);
      out.put('\t');
      // hit_map.c:384-399: rank 0 against max(rank 1, t), later ranks against rank 0
      const int diff = k == 0 ? row[0].score - (row[1].score > a.t ? row[1].score : a.t) : s.score - row[0].score;
      out.num(a.mapq_lut[(diff < -32768 ? -32768 : diff > 32767 ? 32767 : diff) + 32768]);
      out.put('\t');
      const int ed = walk_tb(out, tb, s.tb_len, g, ck, s.ref_start, a.color != 0, false);
      out.lit("\t*\t0\t0\t");
      walk_tb(out, tb, s.tb_len, g, ck, s.ref_start, a.color != 0, true);
      out.put('\t');
      if (qs) {
        if (s.sense == -1) {
          for (int q = L - 1; q >= 0; --q) qc(get2(qs, q));
          if (a.color) qc(a.first_qual ? a.first_qual[i] : 0);
        } else {
          if (a.color) qc(a.first_qual ? a.first_qual[i] : 0);
          for (int q = 0; q < L; ++q) qc(get2(qs, q));
        }
      } else {
        out.put('*');
      }
      out.lit("\tAS:i:");
      out.num(s.score);
      out.lit("\tNM:i:");
      out.num(ed);
      out.lit("\tNH:i:");
      out.num(nrep);
      out.put('\n');
    }
  } else if (a.print_all) {
    name();
    if (a.color) {
      out.lit("\t4\t*\t0\t255\t*\t*\t0\t0\t*\t*\n");
    } else {
      out.lit("\t4\t*\t0\t255\t*\t*\t0\t0\t");
      const uint8_t* cs = a.codes + i * a.code_stride;
      for (int q = 0; q < L; ++q) out.put(BASES[get2(cs, q)]);
      out.put('\t');
      if (qs) for (int q = 0; q < L; ++q) qc(get2(qs, q));
      else out.put('*');
      out.put('\n');
    }
  }
}

}  // namespace

__global__ __launch_bounds__(256) void k_sam_count(SamArgs a) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= a.n) return;
  Sink<false> out{nullptr, 0};
  format_read(a, a.row0 + v, out);
  a.lens[v] = out.n;
  const int64_t i = a.inv ? (int64_t)a.inv[a.row0 + v] : a.row0 + v;
  if (a.mapped && a.slots[i * 4].ref_id != -1) atomicAdd((unsigned long long*)a.mapped, 1ull);
}

__global__ __launch_bounds__(256) void k_sam_write(SamArgs a) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= a.n) return;
  Sink<true> out{a.text + a.offs[v], 0};
  format_read(a, a.row0 + v, out);
}

hipError_t launch_sam(const SamArgs& a, bool write, hipStream_t st) {
  if (a.n == 0) return hipSuccess;
  const dim3 grid((unsigned)((a.n + 255) / 256));
  if (write) hipLaunchKernelGGL(k_sam_write, grid, dim3(256), 0, st, a);
  else hipLaunchKernelGGL(k_sam_count, grid, dim3(256), 0, st, a);
  return hipGetLastError();
}

}  // namespace storm
