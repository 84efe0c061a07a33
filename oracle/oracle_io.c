/*
 * oracle_io.c -- loaders and the -M SAM writer of the CPU restatement (test infrastructure only).
 *
 * Restates reads_data.c, reference_data.c, read_quality.c, load_data.c, codes.c:429-439 and
 * hit_map.c:197-504 / alignment.c:749-887.  Never linked into the product.
 */
#include "storm_oracle.h"

#include <ctype.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_LINE 65536   /* load_data.h:20 MAX_LINE_LEN */

static void *xmalloc(size_t n) {
  void *p = malloc(n ? n : 1);
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); exit(1); }
  return p;
}
static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s ? s : 1);
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); exit(1); }
  return p;
}

static long clen(long len) { return (len >> 2) + ((len & 3) ? 1 : 0); }               /* codes.h:113 */
static long clen_n(long len, int N) { return (clen(len) + (N - 1)) & ~((long)N - 1); }   /* codes.h:116 */
static void set_nth(uint8_t *s, long n, int code) { s[n >> 2] |= (uint8_t)((code & 3) << ((n & 3) << 1)); }

/* read_quality.c:10-16 */
static int quality_level(int q) {
  static const short ub[4] = {2, 7, 10, 93};
  int i = 0;
  while (i < 3 && ub[i] < q) ++i;
  return i;
}
/* read_quality.c:18-29 */
static int parse_qual_char(char c, int base) {
  int v = (int)c - base;
  if (v < 0) { fprintf(stderr, "oracle: quality char below base\n"); exit(-1); }
  if (v > 93) v = 93;
  return v;
}

#define LT_COMMENT 0
#define LT_HEADER 1
#define LT_SEQ 2
#define LT_QHEADER 3
#define LT_QUAL 4
#define LT_UNKNOWN (-1)

/* load_data.c:61-80 (keeps its static state, as the reference does) */
static int line_type(const char *line) {
  static int crt = LT_UNKNOWN;
  int t;
  if (line[0] == '\n' || line[0] == '\r' || line[0] == '\0') return (crt = LT_UNKNOWN);
  else if (crt == LT_COMMENT || line[0] == '#') t = LT_COMMENT;
  else if (crt == LT_HEADER || line[0] == '>') t = LT_HEADER;
  else t = LT_SEQ;
  crt = (t == LT_SEQ || line[strlen(line) - 1] != '\n') ? t : LT_UNKNOWN;
  return t;
}

/* load_data.c:93-117 */
static int line_type_fastq(const char *line, int lineNo, int offset) {
  static int crt = LT_UNKNOWN;
  int t;
  if (line[0] == '\n' || line[0] == '\r' || line[0] == '\0') return LT_UNKNOWN;
  else if (crt == LT_COMMENT || line[0] == '#') t = LT_COMMENT;
  else if ((lineNo - offset) % 4 == 0 && line[0] == '@') t = LT_HEADER;
  else if ((lineNo - offset) % 4 == 2 && line[0] == '+') t = LT_QHEADER;
  else if ((lineNo - offset) % 4 == 1 && crt == LT_HEADER) t = LT_SEQ;
  else if ((lineNo - offset) % 4 == 3 && crt == LT_QHEADER) t = LT_QUAL;
  else t = LT_UNKNOWN;
  crt = (t == LT_SEQ || t == LT_HEADER || t == LT_QUAL || t == LT_QHEADER) ? t : LT_UNKNOWN;
  return t;
}

static void chomp(char *line) {
  size_t n;
  while ((n = strlen(line)) > 0 && (line[n - 1] == '\n' || line[n - 1] == '\r')) line[n - 1] = '\0';
}

/* ------------------------------------------------------------- reference */

static int load_one_reference(FILE *in, int readlen, const orc_params_t *p, orc_ref_t *db) {
  char *line = xmalloc(ORC_LINE);
  long i_start = ftell(in);
  db->name = NULL;
  if (!fgets(line, ORC_LINE, in)) { free(line); return -1; }
  if (line_type(line) == LT_HEADER) {   /* reference_data.c:27-45 */
    int nl = (int)strcspn(line + 1, " \r\n");
    if (nl <= 0) { free(line); return -1; }
    db->name = xmalloc((size_t)nl + 1);
    memcpy(db->name, line + 1, (size_t)nl);
    db->name[nl] = '\0';
    if (db->name[0] == '=' || db->name[0] == '*') db->name[0] = '-';
    i_start = ftell(in);
  } else {
    fseek(in, i_start, SEEK_SET);
  }
  free(line);
  int c;
  long size = 0;
  while ((c = (p->ignore_lowercase ? fgetc(in) : toupper(fgetc(in)))) != EOF) {
    if (isspace(c)) continue;
    if (c == '>') { fseek(in, -1, SEEK_CUR); break; }
    size++;
  }
  if (size == 0) return -1;
  /* reference_data.c:75-88 */
  long shift = clen_n((long)readlen + p->allowed_indels, p->n_bytes);
  long alloc = 2 * shift + clen_n(size, p->n_bytes);
  db->size_shift = shift;
  db->seq_alloc = xcalloc((size_t)alloc, 1);
  db->seq = db->seq_alloc + shift;
  db->mask_alloc = xcalloc((size_t)alloc, 1);
  memset(db->mask_alloc, 0x55, (size_t)shift);
  db->mask = db->mask_alloc + shift;
  fseek(in, i_start, SEEK_SET);
  long i = 0;
  uint8_t code_prev = 0;
  long mstart = -1;
  int cap = 16;
  db->regions = xmalloc(sizeof(orc_region_t) * (size_t)cap);
  db->n_regions = 0;
  db->first_base = 0xff;
  while ((c = (p->ignore_lowercase ? fgetc(in) : toupper(fgetc(in)))) != EOF) {
    uint8_t code;
    if (isspace(c)) continue;
    if (c == '>') { fseek(in, -1, SEEK_CUR); break; }
    if (!(c == 'A' || c == 'C' || c == 'G' || c == 'T')) {   /* reference_data.c:114-123 */
      code = p->color ? code_prev : (uint8_t)(rand() & 3);
      if (mstart == -1) mstart = i;
      set_nth(db->mask, i, 1);
    } else {
      code = (uint8_t)((((c & 0x4) >> 1)) | (((c & 0x4) >> 2) ^ ((c & 0x2) >> 1)));   /* codes.h:39 */
      if (mstart != -1) {
        if (db->n_regions == cap) { cap *= 2; db->regions = realloc(db->regions, sizeof(orc_region_t) * (size_t)cap); }
        db->regions[db->n_regions].start = (int)mstart;
        db->regions[db->n_regions].end = (int)i;
        db->n_regions++;
        mstart = -1;
      }
    }
    if (!p->color) {
      set_nth(db->seq, i, code);
      i++;
    } else {
      if (db->first_base == 0xff) code_prev = db->first_base = code;
      else { set_nth(db->seq, i, code_prev ^ code); i++; code_prev = code; }
    }
  }
  if (mstart != -1) {
    if (db->n_regions == cap) { cap *= 2; db->regions = realloc(db->regions, sizeof(orc_region_t) * (size_t)cap); }
    db->regions[db->n_regions].start = (int)mstart;
    db->regions[db->n_regions].end = (int)i;
    db->n_regions++;
  }
  db->size = (int)i;
  while (i & 3) { set_nth(db->mask, i, 1); i++; }   /* reference_data.c:184-188 */
  memset(db->mask + (i >> 2), 0x55, (size_t)shift);
  return db->size;
}

/* reference_data.c:195-252 */
int orc_load_reference(const char *path, int readlen, const orc_params_t *p, orc_ref_t **refs) {
  FILE *in = fopen(path, "r");
  if (!in) return -1;
  char *line = xmalloc(ORC_LINE);
  int count = 0;
  while (fgets(line, ORC_LINE, in)) if (line_type(line) == LT_HEADER) ++count;
  free(line);
  if (count == 0) count = 1;
  *refs = xcalloc((size_t)count, sizeof(orc_ref_t));
  fseek(in, 0, SEEK_SET);
  int n = 0;
  for (int k = 0; k < count; ++k) {
    if (load_one_reference(in, readlen, p, &(*refs)[k]) <= 0) break;
    n++;
  }
  fclose(in);
  return n;
}

void orc_refs_free(orc_ref_t *refs, int n) {
  if (!refs) return;
  for (int k = 0; k < n; ++k) {
    free(refs[k].name); free(refs[k].seq_alloc); free(refs[k].mask_alloc); free(refs[k].regions);
  }
  free(refs);
}

/* ----------------------------------------------------------------- reads */

/* reads_data.c:485-500: max over prefixes of the last-32-codes window, [atgc] order */
static unsigned long read_tag(const uint8_t *seq, int L) {
  unsigned long tag = 0, c = 0;
  for (int u = 0; u < L; ++u) {
    c <<= 2;
    c |= (unsigned long)((((orc_nth(seq, u) + 1) & 3)) ^ 1);
    if (c > tag) tag = c;
  }
  return tag;
}

/* codes.c:170-197: nucleotide letters through base_to_code (uppercase only), colors by & 3 */
static uint8_t *pack_string(const char *s, int color, int n_bytes) {
  long len = (long)strlen(s);
  long n = clen_n(len, n_bytes);
  uint8_t *out = xcalloc((size_t)n + 8, 1);
  for (long i = 0; i < len; ++i) {
    int code;
    if (color) code = s[i] & 3;
    else code = s[i] == 'A' ? 0 : s[i] == 'C' ? 1 : s[i] == 'G' ? 2 : s[i] == 'T' ? 3 : 0;
    set_nth(out, i, code);
  }
  return out;
}

/* read_quality.c:31-41 */
static uint8_t *compress_quals(const short *q, int len, int shift) {
  uint8_t *out = xcalloc((size_t)clen(len) + 8, 1);
  shift = shift ? 1 : 0;
  for (int i = 0; i < len; ++i) set_nth(out, i, quality_level(q[i + shift]));
  return out;
}

/* reads_data.c:538-564: common printable prefix of the first and last read names */
static char *common_prefix_name(const orc_reads_t *db) {
  if (!(db->n && db->reads[0].info && db->reads[db->n - 1].info)) return NULL;
  const char *a = db->reads[0].info, *b = db->reads[db->n - 1].info;
  int l = 0, t = 0;
  while (a[l] != '\0' && b[l] != '\0' && a[l] == b[l]) {
    if ((t == 0 && ((a[l] >= '!' && a[l] <= ')') || (a[l] >= '+' && a[l] <= '<') || (a[l] >= '>' && a[l] <= '~'))) ||
        (t > 0 && (a[l] >= '!' && a[l] <= '~')))
      t = l + 1;
    l++;
  }
  if (t <= 0) return NULL;
  char *name = xmalloc((size_t)t + 1);
  memcpy(name, a, (size_t)t);
  name[t] = '\0';
  return name;
}

static char *parse_name(char *line) {   /* reads_data.c:460-474 */
  chomp(line);
  int n = 1;
  while (line[n] > ' ' && line[n] <= '~' && line[n] != '@') {
    n++;
    if (n >= 256) break;
  }
  line[n] = '\0';
  char *info = xmalloc((size_t)n);
  memcpy(info, line + 1, (size_t)n - 1);
  info[n - 1] = '\0';
  return info;
}

/* reads_data.c:351-566 */
static long load_fastq(const char *path, const orc_params_t *p, orc_reads_t *db) {
  FILE *in = fopen(path, "r");
  if (!in) return -1;
  char *line = xmalloc(ORC_LINE);
  int lineNo = -1, offset = -1;
  db->n = 0; db->read_len = 0; db->name = NULL;
  long i_start;
  char *rr;
  do {
    offset++; lineNo++;
    i_start = ftell(in);
    rr = fgets(line, ORC_LINE, in);
    if (rr && !db->name && strncmp(line, "# Title: ", 9) == 0) {
      int nl = (int)strcspn(line + 9, "\r\n");
      db->name = xmalloc((size_t)nl + 1);
      memcpy(db->name, line + 9, (size_t)nl);
      db->name[nl] = '\0';
    }
  } while (rr && line_type_fastq(line, lineNo, offset) == LT_COMMENT);
  if (!rr) { fclose(in); free(line); return -1; }
  int lt, next = -1;
  do {
    if ((lt = line_type_fastq(line, lineNo++, offset)) == LT_HEADER && fgets(line, ORC_LINE, in) &&
        (next = line_type_fastq(line, lineNo++, offset)) == LT_SEQ) {
      int len = (int)strcspn(line, "\r\n") - (p->color ? 2 : 0);
      if (db->read_len == 0) {
        db->read_len = len;
        if (len <= 0) { fprintf(stderr, "oracle: corrupt reads\n"); exit(-1); }
      } else if (len != db->read_len) {
        fprintf(stderr, "oracle: read #%ld has length %d, expected %d\n", db->n + 1, len, db->read_len);
        exit(-1);
      }
      ++db->n;
    } else if ((lt == LT_HEADER && next != LT_SEQ) || lt == LT_UNKNOWN) {
      fprintf(stderr, "oracle: corrupt reads file at line %d\n", lineNo);
      exit(-1);
    }
  } while (fgets(line, ORC_LINE, in));
  const int L = db->read_len;
  db->reads = xcalloc((size_t)db->n, sizeof(orc_read_t));
  /* The color FASTQ path of the reference reads quality[read_len], one short past its
   * buffer (reads_data.c:518-529 with compress_quality_sequence shift 1); that value is taken as 0. */
  short *q = xcalloc((size_t)L + 2, sizeof(short));
  fseek(in, i_start, SEEK_SET);
  long i = -1;
  lineNo = offset;
  while (i < db->n) {
    if (!fgets(line, ORC_LINE, in)) {
      if (++i == db->n) continue;
      fclose(in); free(line); free(q);
      return -1;
    }
    switch (line_type_fastq(line, lineNo++, offset)) {
      case LT_HEADER:
        ++i;
        db->reads[i].info = parse_name(line);
        db->reads[i].qual = NULL;
        break;
      case LT_SEQ:
        chomp(line);
        if (!p->color) db->reads[i].seq = pack_string(line, 0, p->n_bytes);
        else {
          uint8_t b = (uint8_t)((((line[0] & 0x4) >> 1)) | (((line[0] & 0x4) >> 2) ^ ((line[0] & 0x2) >> 1)));
          db->reads[i].first_base = (uint8_t)(b ^ (line[1] & 3));
          db->reads[i].seq = pack_string(line + 2, 1, p->n_bytes);
        }
        db->reads[i].tag = read_tag(db->reads[i].seq, L);
        break;
      case LT_QUAL: {
        if (p->color) db->reads[i].first_qual = (uint8_t)quality_level(parse_qual_char(line[0], p->qual_base));
        chomp(line);
        memset(q, 0, sizeof(short) * ((size_t)L + 2));
        int lim = (int)strlen(line) < L ? (int)strlen(line) : L;
        for (int k = p->color ? 1 : 0; k < lim; ++k) q[k] = (short)parse_qual_char(line[k], p->qual_base);
        db->reads[i].qual = compress_quals(q, L, p->color ? 1 : 0);
        break;
      }
      default:
        break;
    }
  }
  fclose(in);
  free(line); free(q);
  if (!db->name) db->name = common_prefix_name(db);
  return db->n;
}

/* reads_data.c:73-349 */
static long load_fasta(const char *path, const char *qpath, const orc_params_t *p, orc_reads_t *db) {
  FILE *in = fopen(path, "r");
  if (!in) return -1;
  FILE *qin = NULL;
  if (qpath) { qin = fopen(qpath, "r"); if (!qin) { fclose(in); return -1; } }
  char *line = xmalloc(ORC_LINE);
  db->n = 0; db->read_len = 0; db->name = NULL;
  long i_start;
  char *rr;
  do {
    i_start = ftell(in);
    rr = fgets(line, ORC_LINE, in);
    if (rr && !db->name && strncmp(line, "# Title: ", 9) == 0) {
      int nl = (int)strcspn(line + 9, "\r\n");
      db->name = xmalloc((size_t)nl + 1);
      memcpy(db->name, line + 9, (size_t)nl);
      db->name[nl] = '\0';
    }
  } while (rr && line_type(line) == LT_COMMENT);
  if (!rr) { fclose(in); free(line); return -1; }
  do {
    if (line_type(line) == LT_HEADER && fgets(line, ORC_LINE, in) && line_type(line) == LT_SEQ) {
      int len = (int)strcspn(line, "\r\n") - (p->color ? 2 : 0);
      while ((rr = fgets(line, ORC_LINE, in)) && line_type(line) == LT_SEQ) len += (int)strcspn(line, "\r\n");
      if (db->read_len == 0) {
        db->read_len = len;
        if (len <= 0) { fprintf(stderr, "oracle: corrupt reads\n"); exit(-1); }
      } else if (len != db->read_len) {
        fprintf(stderr, "oracle: read #%ld has length %d, expected %d\n", db->n + 1, len, db->read_len);
        exit(-1);
      }
      ++db->n;
    }
  } while (rr);
  const int L = db->read_len;
  db->reads = xcalloc((size_t)db->n, sizeof(orc_read_t));
  fseek(in, i_start, SEEK_SET);
  char *buf = xcalloc((size_t)L + 3, 1);
  long i = 0;
  int done = 0;
  rr = fgets(line, ORC_LINE, in);
  while (rr && !done) {
    int t = line_type(line);
    if (t == LT_HEADER) {
      db->reads[i].info = parse_name(line);
      db->reads[i].qual = NULL;
      rr = fgets(line, ORC_LINE, in);
    } else if (t == LT_SEQ) {
      char *pos = buf;
      int remain = L + (p->color ? 2 : 0);
      memset(buf, 0, (size_t)L + 3);
      do {
        chomp(line);
        int ll = (int)strlen(line);
        if (ll <= remain) { memcpy(pos, line, (size_t)ll); pos += ll; remain -= ll; }
        else { fprintf(stderr, "oracle: read #%ld has an unexpected format\n", i + 1); exit(-1); }
      } while ((rr = fgets(line, ORC_LINE, in)) && line_type(line) == LT_SEQ);
      if (!p->color) db->reads[i].seq = pack_string(buf, 0, p->n_bytes);
      else {
        uint8_t b = (uint8_t)((((buf[0] & 0x4) >> 1)) | (((buf[0] & 0x4) >> 2) ^ ((buf[0] & 0x2) >> 1)));
        db->reads[i].first_base = (uint8_t)(b ^ (buf[1] & 3));
        db->reads[i].seq = pack_string(buf + 2, 1, p->n_bytes);
      }
      db->reads[i].tag = read_tag(db->reads[i].seq, L);
      i++;
      if (i == db->n) done = 1;
    } else {
      rr = fgets(line, ORC_LINE, in);
    }
  }
  free(buf);
  fclose(in);
  if (qin) {   /* reads_data.c:297-326 */
    short *q = xcalloc((size_t)L + 2, sizeof(short));
    int want = L + (p->color ? 1 : 0);
    for (long k = 0; k < db->n;) {
      if (!fgets(line, ORC_LINE, qin)) break;
      if (line_type(line) == LT_SEQ) {
        char *copy = strdup(line), *tok, *s = copy;
        for (int j = 0; j < want; ++j, s = NULL) {
          tok = strtok(s, " ");
          if (!tok) break;
          q[j] = (short)atoi(tok);
        }
        free(copy);
        if (p->color) {
          db->reads[k].first_qual = (uint8_t)quality_level(q[0]);
          db->reads[k].qual = compress_quals(q, L, 1);
        } else {
          db->reads[k].qual = compress_quals(q, L, 0);
        }
        k++;
      }
    }
    free(q);
    fclose(qin);
  }
  free(line);
  if (!db->name) db->name = common_prefix_name(db);
  return db->n;
}

/* reads_data.c:568-578: dispatch on the file extension */
long orc_load_reads(const char *path, const char *qual_path, const orc_params_t *p, orc_reads_t *db) {
  memset(db, 0, sizeof(*db));
  const char *dot = strrchr(path, '.');
  if (dot && (!strcmp(dot, ".fastq") || !strcmp(dot, ".fq") || !strcmp(dot, ".fastq2") || !strcmp(dot, ".fq2")))
    return load_fastq(path, p, db);
  return load_fasta(path, qual_path, p, db);
}

/* reads_data.c:41-67: qsort by tag; glibc 2.35 sorts these arrays with its merge sort,
 * which keeps equal tags in input order, so a stable merge sort is restated here. */
void orc_sort_reads(orc_reads_t *db) {
  long n = db->n;
  if (n < 2) return;
  orc_read_t *tmp = xmalloc(sizeof(orc_read_t) * (size_t)n);
  orc_read_t *a = db->reads;
  for (long width = 1; width < n; width *= 2) {
    for (long lo = 0; lo < n; lo += 2 * width) {
      long mid = lo + width < n ? lo + width : n, hi = lo + 2 * width < n ? lo + 2 * width : n;
      long x = lo, y = mid, k = lo;
      while (x < mid && y < hi) tmp[k++] = (a[y].tag < a[x].tag) ? a[y++] : a[x++];
      while (x < mid) tmp[k++] = a[x++];
      while (y < hi) tmp[k++] = a[y++];
    }
    memcpy(a, tmp, sizeof(orc_read_t) * (size_t)n);
  }
  free(tmp);
}

void orc_reads_free(orc_reads_t *db) {
  for (long i = 0; i < db->n; ++i) { free(db->reads[i].info); free(db->reads[i].seq); free(db->reads[i].qual); }
  free(db->reads);
  free(db->name);
  memset(db, 0, sizeof(*db));
}

/* reads_data.c:9-35, codes.c:429-439, read_quality.c:47-64.  dst buffers are zeroed by the caller. */
void orc_read_reverse(const orc_read_t *src, orc_read_t *dst, int len, int color) {
  if (color) {
    uint8_t comp = 0;
    for (int k = 0; k < len; ++k) comp ^= (uint8_t)orc_nth(src->seq, k);
    dst->first_base = (uint8_t)((~(src->first_base ^ comp)) & 3);
  }
  for (int k = 0; k < len; ++k) {
    int c = orc_nth(src->seq, k);
    set_nth(dst->seq, len - k - 1, color ? c : ((~c) & 3));
  }
  if (src->qual) {
    if (!color) {
      for (int k = 0; k < len; ++k) set_nth(dst->qual, k, orc_nth(src->qual, len - k - 1));
    } else {
      set_nth(dst->qual, len - 1, src->first_qual);
      for (int k = 0; k < len - 1; ++k) set_nth(dst->qual, k, orc_nth(src->qual, len - k - 2));
      dst->first_qual = (uint8_t)orc_nth(src->qual, len - 1);
    }
  }
}

/* ------------------------------------------------------------------- SAM */

/* alignment.c:749-887 */
int orc_traceback_to_cigar(const uint8_t *tb, int tb_len, const uint8_t *ref_seq, long ref_start,
                           int color, uint8_t ref_start_symbol, char *cigar, char *seq) {
  static const char BASES[4] = {'A', 'C', 'G', 'T'};
  int ed = 0, span = color ? 1 : 0, slen = 0;
  int prev = 0x3;
  char *cp = cigar;
  long ref_idx = ref_start;
  uint8_t ref_hist = 0, read_hist = 0;
  uint8_t *codes = xmalloc((size_t)tb_len + 4);
  cigar[0] = '\0';
  if (color) codes[slen++] = 0;
  for (int i = 0; i < tb_len; ++i) {
    int cell = (tb[i >> 1] >> ((i & 1) ? 0 : 4)) & 0xF;
    int type;
    if ((cell & 0xC) == 0x4) type = 0x4;
    else if (color && (cell & 0xC) == 0x8) type = 0x8;
    else if ((cell & 0xC) == 0xC) type = 0xC;
    else type = cell;
    uint8_t code = (uint8_t)(cell & 3);
    switch (type) {
      case 0x3:
        ed--;
        if (!color) code = (uint8_t)orc_nth(ref_seq, ref_idx);
        /* fall through */
      case 0x4:
        ed++;
        if (color) {
          ref_hist ^= (uint8_t)orc_nth(ref_seq, ref_idx);
          code = (uint8_t)(ref_hist ^ read_hist);
        }
        codes[slen++] = code;
        if (color) read_hist ^= code;
        ref_idx++;
        break;
      case 0x8:
        ed++;
        codes[slen++] = code;
        read_hist ^= code;
        ref_hist ^= (uint8_t)orc_nth(ref_seq, ref_idx);
        ref_idx++;
        break;
      case 0xC:
        ed++;
        codes[slen++] = code;
        if (color) read_hist ^= code;
        break;
      case 0x1:
        ed++;
        if (color) ref_hist ^= (uint8_t)orc_nth(ref_seq, ref_idx);
        ref_idx++;
        break;
      default:
        break;
    }
    int crt = (type == 0x1 || type == 0xC) ? type : 0x3;
    if (crt == prev) ++span;
    else {
      if (span) cp += sprintf(cp, "%d%c", span, prev == 0x3 ? 'M' : prev == 0xC ? 'I' : prev == 0x1 ? 'D' : '?');
      span = 1;
      prev = crt;
    }
  }
  if (span) sprintf(cp, "%d%c", span, prev == 0x3 ? 'M' : prev == 0xC ? 'I' : prev == 0x1 ? 'D' : '?');
  if (!color) {
    for (int u = 0; u < slen; ++u) seq[u] = BASES[codes[u]];
  } else {
    uint8_t c = ref_start_symbol;
    for (int u = 0; u < slen; ++u) { c = (uint8_t)(c ^ codes[u]); seq[u] = BASES[c & 3]; }
  }
  seq[slen] = '\0';
  free(codes);
  return ed;
}

static int row_reportable(const orc_hit_t *row, int rank, const orc_params_t *p) {   /* hit_map.c:359-370 */
  int M = p->unordered_M;
  return rank < M && row[rank].ref_start != ORC_UNMAPPED &&
         ((p->uniqueness == 0) ||
          ((rank + 1 < M) && ((row[rank].score - row[rank + 1].score) >= p->uniqueness_threshold)) ||
          (rank + 1 == M));
}

/* MAPQ of rank k of a hit map row (hit_map.c:374-399, genome_map.c:700-724) */
static int orc_mapq(const orc_hit_t *row, int k, const orc_params_t *p) {
  const double lg = log(.3);
  const int delta_m = p->match - p->mismatch;   /* main.c:435 */
  int s1 = row[0].score;
  int s2 = (k == 0) ? (row[1].score > p->t ? row[1].score : p->t) : row[k].score;
  double n = (k == 0) ? (double)(s1 - s2) / (double)delta_m : (double)(s2 - s1) / (double)delta_m;
  int mq = -4.343 * (n * lg);
  mq = (int)(mq - 4.343 * lg);
  mq = mq >= 0 ? mq : 0;
  return mq < 254 ? mq : 254;
}

/* base-space symbol before each position of a color reference (hit_map.c:236-249; the ordered
 * writer's running ref_symbol, genome_map.c:679,789 is the same sequence) */
static uint8_t **orc_ref_symbols(const orc_ref_t *refs, int n_refs) {
  uint8_t **sym = xcalloc((size_t)n_refs, sizeof(uint8_t *));
  for (int r = 0; r < n_refs; ++r) {
    sym[r] = xcalloc((size_t)clen((long)refs[r].size + 1) + 1, 1);
    uint8_t s = refs[r].first_base;
    long k;
    for (k = 0; k < refs[r].size; ++k) { set_nth(sym[r], k, s); s = (uint8_t)(s ^ orc_nth(refs[r].seq, k)); }
    set_nth(sym[r], k, s);
  }
  return sym;
}

static void orc_sam_header(FILE *out, const char *sort_order, const orc_reads_t *db, const orc_ref_t *refs,
                           int n_refs, const orc_params_t *p, int argc, char **argv) {
  fprintf(out, "@HD\tVN:1.4\tSO:%s\n", sort_order);
  for (int r = 0; r < n_refs; ++r)
    fprintf(out, "@SQ\tSN:%s\tLN:%ld\n", refs[r].name ? refs[r].name : "*", (long)(refs[r].size + (p->color ? 1 : 0)));
  fprintf(out, "@RG\tID:%s\tSM:%s\n", db->name ? db->name : "*", db->name ? db->name : "*");
  fprintf(out, "@PG\tID:%s\tVN:%s\tCL:\"", "SToRM", "0.0101");
  for (int a = 0; a < argc; ++a) fprintf(out, "%s ", argv[a]);
  fprintf(out, "\"\n");
}

/* the fields of one mapped record from FLAG to QUAL (hit_map.c:401-448, genome_map.c:727-771) */
static void orc_sam_record(FILE *out, const orc_read_t *rd, const orc_hit_t *h, int flag, int aq, const orc_ref_t *refs,
                           uint8_t **sym, const orc_params_t *p, int L, char *cigar, char *seq) {
  static const short UB[4] = {2, 7, 10, 93};
  int rid = h->ref_id;
  long rs = h->ref_start;
  int ed = orc_traceback_to_cigar(h->traceback, h->traceback_len, refs[rid].seq, rs, p->color,
                                  p->color ? (uint8_t)orc_nth(sym[rid], rs) : 0, cigar, seq);
  fprintf(out, "%s\t%d\t%s\t%ld\t%d\t%s\t", rd->info, flag, refs[rid].name ? refs[rid].name : "*", rs + 1, aq, cigar);
  fprintf(out, "*\t0\t0\t");
  fprintf(out, "%s\t", seq);
  if (rd->qual) {
    if (h->sense & ORC_SENSE_REV) {
      for (int q = L - 1; q >= 0; --q) {
        int v = p->qual_base + UB[orc_nth(rd->qual, q)];
        fputc(v < 126 ? v : 126, out);
      }
      if (p->color) { int v = p->qual_base + UB[rd->first_qual]; fputc(v < 126 ? v : 126, out); }
    } else {
      if (p->color) { int v = p->qual_base + UB[rd->first_qual]; fputc(v < 126 ? v : 126, out); }
      for (int q = 0; q < L; ++q) {
        int v = p->qual_base + UB[orc_nth(rd->qual, q)];
        fputc(v < 126 ? v : 126, out);
      }
    }
  } else {
    fputc('*', out);
  }
  fprintf(out, "\tAS:i:%d\tNM:i:%d", h->score, ed);
}

/* hit_map.c:197-504 */
void orc_write_sam(FILE *out, const orc_result_t *res, const orc_reads_t *db, const orc_ref_t *refs, int n_refs,
                   const orc_params_t *p, int argc, char **argv) {
  static const char BASES[4] = {'A', 'C', 'G', 'T'};
  static const short UB[4] = {2, 7, 10, 93};
  const int L = db->read_len;
  uint8_t **sym = p->color ? orc_ref_symbols(refs, n_refs) : NULL;
  orc_sam_header(out, "unsorted", db, refs, n_refs, p, argc, argv);
  char *cigar = xmalloc((size_t)L * 4 + 64);
  char *seq = xmalloc((size_t)L + 64);
  for (long id = 0; id < res->n_reads; ++id) {
    const orc_hit_t *row = res->map[id];
    const orc_read_t *rd = &db->reads[id];
    if (row) {
      int nrep = 0;
      for (int k = 0; row_reportable(row, k, p); ++k) nrep++;
      for (int k = 0; row_reportable(row, k, p); ++k) {
        int flag = 0;
        if (row[k].sense & ORC_SENSE_REV) flag |= 0x10;
        if (k) flag |= 0x100;
        orc_sam_record(out, rd, &row[k], flag, orc_mapq(row, k, p), refs, sym, p, L, cigar, seq);
        fprintf(out, "\tNH:i:%d\n", nrep);
      }
    } else if (p->print_all) {
      if (p->color) {
        fprintf(out, "%s\t4\t*\t0\t255\t*\t*\t0\t0\t*\t*\n", rd->info);
      } else {
        fprintf(out, "%s\t4\t*\t0\t255\t*\t*\t0\t0\t", rd->info);
        for (int q = 0; q < L; ++q) fputc(BASES[orc_nth(rd->seq, q)], out);
        fputc('\t', out);
        if (rd->qual) {
          for (int q = 0; q < L; ++q) { int v = p->qual_base + UB[orc_nth(rd->qual, q)]; fputc(v < 126 ? v : 126, out); }
        } else fputc('*', out);
        fputc('\n', out);
      }
    }
  }
  fflush(out);
  free(cigar); free(seq);
  if (sym) { for (int r = 0; r < n_refs; ++r) free(sym[r]); free(sym); }
}

/* hit_map__generate_unmapped_FASTQ_output (hit_map.c:514-578): the reads without any hit */
void orc_write_unmapped_fastq(FILE *out, const orc_result_t *res, const orc_reads_t *db, const orc_params_t *p) {
  static const char BASES[4] = {'A', 'C', 'G', 'T'};
  static const short UB[4] = {2, 7, 10, 93};
  const int L = db->read_len;
  for (long id = 0; id < res->n_reads; ++id) {
    if (res->map[id]) continue;
    const orc_read_t *rd = &db->reads[id];
    if (rd->info) fprintf(out, "@%s\n", rd->info); else fprintf(out, "@%ld\n", id);
    if (p->color) {
      fputc(BASES[rd->first_base & 3], out);
      for (int k = 0; k < L; ++k) fputc('0' + orc_nth(rd->seq, k), out);
    } else {
      for (int k = 0; k < L; ++k) fputc(BASES[orc_nth(rd->seq, k)], out);
    }
    fputc('\n', out);
    if (rd->info) fprintf(out, "+%s\n", rd->info); else fprintf(out, "+%ld\n", id);
    if (rd->qual) {
      if (p->color) { int v = p->qual_base + UB[rd->first_qual]; fputc(v < 126 ? v : 126, out); }
      for (int q = 0; q < L; ++q) { int v = p->qual_base + UB[orc_nth(rd->qual, q)]; fputc(v < 126 ? v : 126, out); }
    } else {
      const int v = p->qual_base + 93 < 126 ? p->qual_base + 93 : 126;   /* MAX_READ_QUALITY */
      for (int q = 0; q < L + (p->color ? 1 : 0); ++q) fputc(v, out);
    }
    fputc('\n', out);
  }
  fflush(out);
}

/* ------------------------------------------------------------ ordered mode (genome_map.c) */

#define GM_NO_CODE 4   /* GENOME_MAP_NO_CODE = CODE_COUNT (genome_map.c:26) */

/* per mapped position (GenomeMapPositionType, genome_map.h:31-48, COMPRESSED_COUNTERS): the reads
 * placed here in placement order, and 6-bit counters of codes 0..3 and the gap (4) packed in an
 * unsigned, for the aligned code and for each insertion offset 0..I-1 */
typedef struct {
  int *read, *rank;
  int n;
  unsigned aligned;
  unsigned *inserted;
} orc_gpos_t;

/* COUNTER_GET / COUNTER_INC (genome_map.c:38-47): the increment wraps at 64 */
static unsigned gm_get(unsigned c, int code) { return (c >> (code * 6)) & 0x3Fu; }
static void gm_inc(unsigned *c, int code) {
  unsigned u = (gm_get(*c, code) + 1) & 0x3Fu;
  *c = (*c & ~(0x3Fu << (code * 6))) | (u << (code * 6));
}

/* genome_map__create_position_data (genome_map.c:68-99); positions 0..size exist (GENOME_MAP_SIZE) */
static orc_gpos_t *gm_at(orc_gpos_t **g, long size, long pos, int indels, int greedy) {
  if (pos < 0 || pos > size) {
    fprintf(stderr, "oracle: genome map position %ld outside 0..%ld\n", pos, size);
    exit(2);
  }
  if (!g[pos]) {
    g[pos] = xcalloc(1, sizeof(orc_gpos_t));
    if (!greedy && indels) g[pos]->inserted = xcalloc((size_t)indels, sizeof(unsigned));
  }
  return g[pos];
}

/* genome_map__set_read_idx_at (genome_map.c:123-138) */
static void gm_place(orc_gpos_t **g, long size, long pos, int read, int rank, int indels, int greedy) {
  orc_gpos_t *x = gm_at(g, size, pos, indels, greedy);
  x->read = realloc(x->read, sizeof(int) * (size_t)(x->n + 1));
  x->rank = realloc(x->rank, sizeof(int) * (size_t)(x->n + 1));
  x->read[x->n] = read;
  x->rank[x->n] = rank;
  x->n++;
}

/* genome_map__get_direct_consensus_code (genome_map.c:308-330) */
static int gm_consensus(orc_gpos_t **g, long size, long pos, int offset, int reference_code) {
  if (pos < 0 || pos > size) {
    fprintf(stderr, "oracle: genome map position %ld outside 0..%ld\n", pos, size);
    exit(2);
  }
  if (!g[pos]) return reference_code;
  unsigned c = offset == -1 ? g[pos]->aligned : g[pos]->inserted[offset];
  int best = reference_code;
  for (int i = 0; i <= GM_NO_CODE; ++i)
    if (gm_get(c, i) > gm_get(c, best)) best = i;
  if (gm_get(c, best) == 0) best = GM_NO_CODE;
  return best;
}

/* genome_map__update_consensus_code (genome_map.c:340-349) */
static void gm_count(orc_gpos_t **g, long size, long pos, int offset, int code, int indels) {
  orc_gpos_t *x = gm_at(g, size, pos, indels, 0);
  if (offset >= 0) gm_inc(&x->inserted[offset], code);
  else gm_inc(&x->aligned, code);
}

static int tb_fwd(const uint8_t *tb, int i) { return (tb[i >> 1] >> ((i & 1) ? 0 : 4)) & 0xF; }   /* alignment.h:82 */

/* GET_PAIR_TYPE (alignment.h:87-94) */
static int pair_type(int cell, int color) {
  if ((cell & 0xC) == 0x4) return 0x4;
  if (color && (cell & 0xC) == 0x8) return 0x8;
  if ((cell & 0xC) == 0xC) return 0xC;
  return cell;
}

/* genome_map__compute_adjusted_score (genome_map.c:359-425) */
static int gm_adjusted(orc_gpos_t **g, const orc_hit_t *h, const orc_ref_t *ref, const orc_params_t *p) {
  if (h->ref_start == ORC_UNMAPPED) return ORC_MIN_SCORE;
  int score = 0, indel = 0;
  long pos = h->ref_start;
  for (int i = 0; i < h->traceback_len; ++i) {
    int cell = tb_fwd(h->traceback, i);
    switch (pair_type(cell, p->color)) {
      case 0x3: {
        int rc = orc_nth(ref->seq, pos);
        score += (rc == gm_consensus(g, ref->size, pos, -1, rc)) ? p->match : p->mismatch;
        ++pos; indel = 0;
      } break;
      case 0x1:
        score += (gm_consensus(g, ref->size, pos, -1, GM_NO_CODE) == GM_NO_CODE) ? p->match : p->mismatch;
        ++pos; indel = 0;
        break;
      case 0x4: case 0x8: {
        int c = cell & 3;
        score += (c == gm_consensus(g, ref->size, pos, -1, c)) ? p->match : p->mismatch;
        ++pos; indel = 0;
      } break;
      case 0xC: {
        int c = cell & 3;
        score += (c == gm_consensus(g, ref->size, pos, indel, c)) ? p->match : p->mismatch;
        ++indel;
      } break;
      default: break;
    }
  }
  return score;
}

/* genome_map__radix_sort_reads (genome_map.c:252-298): LSD radix over the 8 hex digits of the
 * int score, each digit's bucket reversed, so the order is descending in the unsigned 32-bit
 * value of the score, ties in the incoming order */
static void gm_sort(int *ids, long n, orc_hit_t **map) {
  int *tmp = xmalloc(sizeof(int) * (size_t)(n ? n : 1));
  for (int d = 0; d < 8; ++d) {
    long count[16] = {0}, index[16];
    for (long j = 0; j < n; ++j) count[15 - (((unsigned)map[ids[j]][0].score >> (4 * d)) & 15u)]++;
    index[0] = 0;
    for (int b = 1; b < 16; ++b) index[b] = index[b - 1] + count[b - 1];
    for (long j = 0; j < n; ++j) tmp[index[15 - (((unsigned)map[ids[j]][0].score >> (4 * d)) & 15u)]++] = ids[j];
    memcpy(ids, tmp, sizeof(int) * (size_t)n);
  }
  free(tmp);
}

void orc_write_sam_ordered(FILE *out, const orc_result_t *res, const orc_reads_t *db, const orc_ref_t *refs,
                           int n_refs, const orc_params_t *p, int argc, char **argv) {
  const int I = p->allowed_indels;   /* hit_map__create(size, allowed_indels), run.c:818 */
  orc_gpos_t ***g = xcalloc((size_t)n_refs, sizeof(orc_gpos_t **));
  for (int r = 0; r < n_refs; ++r) g[r] = xcalloc((size_t)refs[r].size + 1, sizeof(orc_gpos_t *));
  if (p->greedy) {   /* genome_map__build_greedy / __map_read_greedy (genome_map.c:186-232) */
    for (long id = 0; id < res->n_reads; ++id) {
      const orc_hit_t *row = res->map[id];
      if (!row || row[0].ref_start == ORC_UNMAPPED) continue;
      if (p->uniqueness == 0 || row[0].score - row[1].score >= p->uniqueness_threshold)
        gm_place(g[row[0].ref_id], refs[row[0].ref_id].size, row[0].ref_start, (int)id, 0, I, 1);
    }
  } else {   /* genome_map__build_contextual (genome_map.c:432-536) */
    int *ids = xmalloc(sizeof(int) * (size_t)(res->n_reads ? res->n_reads : 1));
    long nm = 0;
    for (long id = 0; id < res->n_reads; ++id)
      if (res->map[id]) ids[nm++] = (int)id;
    gm_sort(ids, nm, res->map);
    for (long s = 0; s < nm; ++s) {
      const int id = ids[s];
      const orc_hit_t *row = res->map[id];
      int adj[ORC_MAP_DETAIL];
      for (int k = 0; k < ORC_MAP_DETAIL; ++k) adj[k] = ORC_MIN_SCORE;
      int best_rank = ORC_UNMAPPED, best = ORC_MIN_SCORE;
      for (int k = 0; k < ORC_MAP_DETAIL && row[k].ref_start != ORC_UNMAPPED; ++k) {
        adj[k] = gm_adjusted(g[row[k].ref_id], &row[k], &refs[row[k].ref_id], p);
        if (adj[k] > best) { best_rank = k; best = adj[k]; }
      }
      if (best_rank == ORC_UNMAPPED || !(p->uniqueness == 0 || adj[0] - adj[1] > p->uniqueness_threshold)) continue;
      const orc_hit_t *h = &row[best_rank];
      orc_gpos_t **gr = g[h->ref_id];
      const orc_ref_t *ref = &refs[h->ref_id];
      long pos = h->ref_start;
      gm_place(gr, ref->size, pos, id, best_rank, I, 0);
      int ins = -1;
      for (int k = 0; k < h->traceback_len; ++k) {
        int cell = tb_fwd(h->traceback, k);
        switch (pair_type(cell, p->color)) {
          case 0xC: ins++; gm_count(gr, ref->size, pos, ins, cell & 3, I); break;
          case 0x3: ins = -1; gm_count(gr, ref->size, pos, -1, orc_nth(ref->seq, pos), I); break;
          case 0x4: case 0x8: ins = -1; gm_count(gr, ref->size, pos, -1, cell & 3, I); break;
          case 0x1: ins = -1; gm_count(gr, ref->size, pos, -1, GM_NO_CODE, I); break;
          default: break;
        }
        for (int q = ins + 1; q < I; ++q) gm_count(gr, ref->size, pos, q, GM_NO_CODE, I);
        if (ins < 0) pos++;
      }
    }
    free(ids);
  }
  /* genome_map__generate_SAM_output (genome_map.c:551-800) */
  const int L = db->read_len;
  uint8_t **sym = p->color ? orc_ref_symbols(refs, n_refs) : NULL;
  orc_sam_header(out, "coordinate", db, refs, n_refs, p, argc, argv);
  char *cigar = xmalloc((size_t)L * 4 + 64);
  char *seq = xmalloc((size_t)L + 64);
  for (int r = 0; r < n_refs; ++r) {
    for (long pos = 0; pos <= refs[r].size; ++pos) {
      orc_gpos_t *x = g[r][pos];
      if (!x) continue;
      for (int k = 0; k < x->n; ++k) {
        const orc_hit_t *row = res->map[x->read[k]];
        const orc_hit_t *h = &row[x->rank[k]];
        int flag = (h->sense & ORC_SENSE_REV) ? 0x10 : 0;
        orc_sam_record(out, &db->reads[x->read[k]], h, flag, orc_mapq(row, x->rank[k], p), refs, sym, p, L, cigar, seq);
        fputc('\n', out);
      }
      free(x->read); free(x->rank); free(x->inserted); free(x);
    }
    free(g[r]);
  }
  free(g);
  fflush(out);
  free(cigar); free(seq);
  if (sym) { for (int r = 0; r < n_refs; ++r) free(sym[r]); free(sym); }
}

/* ----------------------------------------------------------------- test helpers */

/* Sequence ref_no of a FASTA file as load_reference holds it: a malloc'd copy of the packed
 * codes (size codes, 4 per byte, no padding) and of the masked regions [start, end).  Both
 * helpers load the file from rand()'s initial state (srand(1)), as a fresh process does, so the
 * codes drawn for N are the same in both. */
int orc_fasta_sequence(const char *path, int color, int ignore_lowercase, int ref_no, uint8_t **packed, int *size,
                       int **starts, int **ends, int *n_regions) {
  orc_params_t p;
  orc_params_default(&p, color);
  p.ignore_lowercase = ignore_lowercase;
  orc_ref_t *refs = NULL;
  srand(1);
  int n = orc_load_reference(path, 100, &p, &refs);
  if (n <= ref_no || ref_no < 0) { orc_refs_free(refs, n > 0 ? n : 0); return -1; }
  const orc_ref_t *r = &refs[ref_no];
  const size_t bytes = ((size_t)r->size + 3) / 4;
  *packed = xmalloc(bytes + 1);
  memcpy(*packed, r->seq, bytes);
  *size = r->size;
  *n_regions = r->n_regions;
  *starts = xmalloc(sizeof(int) * (size_t)(r->n_regions + 1));
  *ends = xmalloc(sizeof(int) * (size_t)(r->n_regions + 1));
  for (int k = 0; k < r->n_regions; ++k) { (*starts)[k] = r->regions[k].start; (*ends)[k] = r->regions[k].end; }
  orc_refs_free(refs, n);
  return 0;
}

/* The index of one seed pattern over sequence ref_no of a FASTA file (index.c:19-195):
 * malloc'd first[4^w + 1] and map[first[4^w]]. */
int orc_fasta_index(const char *path, int color, int ignore_lowercase, int ref_no, const char *seed,
                    int stdev_distance, int **first, int *n_first, int **map, int *n_map) {
  orc_params_t p;
  orc_params_default(&p, color);
  p.ignore_lowercase = ignore_lowercase;
  p.filter_repeats = stdev_distance > 0;
  p.stdev_distance = stdev_distance > 0 ? stdev_distance : 0;
  orc_seed_t *seeds = NULL;
  int ns = orc_seed_parse_list(seed, &seeds);
  if (ns != 1) { if (ns > 0) orc_seeds_free(seeds, ns); return -1; }
  orc_ref_t *refs = NULL;
  srand(1);
  int n = orc_load_reference(path, 100, &p, &refs);
  if (n <= ref_no || ref_no < 0) { orc_refs_free(refs, n > 0 ? n : 0); orc_seeds_free(seeds, ns); return -1; }
  orc_index_t idx;
  orc_index_build(&refs[ref_no], &seeds[0], &p, &idx);
  *n_first = idx.index_size + 1;
  *first = xmalloc(sizeof(int) * (size_t)*n_first);
  memcpy(*first, idx.first, sizeof(int) * (size_t)*n_first);
  *n_map = idx.first[idx.index_size];
  *map = xmalloc(sizeof(int) * (size_t)(*n_map + 1));
  memcpy(*map, idx.map, sizeof(int) * (size_t)*n_map);
  orc_index_free(&idx);
  orc_refs_free(refs, n);
  orc_seeds_free(seeds, ns);
  return 0;
}

void orc_free(void *p) { free(p); }
