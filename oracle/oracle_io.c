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

/* hit_map.c:197-504 */
void orc_write_sam(FILE *out, const orc_result_t *res, const orc_reads_t *db, const orc_ref_t *refs, int n_refs,
                   const orc_params_t *p, int argc, char **argv) {
  static const char BASES[4] = {'A', 'C', 'G', 'T'};
  static const short UB[4] = {2, 7, 10, 93};
  const int L = db->read_len;
  uint8_t **sym = NULL;
  if (p->color) {   /* hit_map.c:236-249: base-space symbols of each color reference */
    sym = xcalloc((size_t)n_refs, sizeof(uint8_t *));
    for (int r = 0; r < n_refs; ++r) {
      sym[r] = xcalloc((size_t)clen((long)refs[r].size + 1) + 1, 1);
      uint8_t s = refs[r].first_base;
      long k;
      for (k = 0; k < refs[r].size; ++k) { set_nth(sym[r], k, s); s = (uint8_t)(s ^ orc_nth(refs[r].seq, k)); }
      set_nth(sym[r], k, s);
    }
  }
  fprintf(out, "@HD\tVN:1.4\tSO:unsorted\n");
  for (int r = 0; r < n_refs; ++r)
    fprintf(out, "@SQ\tSN:%s\tLN:%ld\n", refs[r].name ? refs[r].name : "*", (long)(refs[r].size + (p->color ? 1 : 0)));
  fprintf(out, "@RG\tID:%s\tSM:%s\n", db->name ? db->name : "*", db->name ? db->name : "*");
  fprintf(out, "@PG\tID:%s\tVN:%s\tCL:\"", "SToRM", "0.0101");
  for (int a = 0; a < argc; ++a) fprintf(out, "%s ", argv[a]);
  fprintf(out, "\"\n");
  char *cigar = xmalloc((size_t)L * 4 + 64);
  char *seq = xmalloc((size_t)L + 64);
  const double lg = log(.3);
  const int delta_m = p->match - p->mismatch;
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
        int aq;
        {
          int s1 = row[0].score;
          int s2 = (k == 0) ? (row[1].score > p->t ? row[1].score : p->t) : row[k].score;
          double n = (k == 0) ? (double)(s1 - s2) / (double)delta_m : (double)(s2 - s1) / (double)delta_m;
          int mq = -4.343 * (n * lg);
          mq = (int)(mq - 4.343 * lg);
          mq = mq >= 0 ? mq : 0;
          mq = mq < 254 ? mq : 254;
          aq = mq;
        }
        int rid = row[k].ref_id;
        long rs = row[k].ref_start;
        int ed = orc_traceback_to_cigar(row[k].traceback, row[k].traceback_len, refs[rid].seq, rs, p->color,
                                        p->color ? (uint8_t)orc_nth(sym[rid], rs) : 0, cigar, seq);
        fprintf(out, "%s\t%d\t%s\t%ld\t%d\t%s\t", rd->info, flag, refs[rid].name ? refs[rid].name : "*", rs + 1, aq, cigar);
        fprintf(out, "*\t0\t0\t");
        fprintf(out, "%s\t", seq);
        if (rd->qual) {
          if (row[k].sense & ORC_SENSE_REV) {
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
        fprintf(out, "\tAS:i:%d\tNM:i:%d\tNH:i:%d\n", row[k].score, ed, nrep);
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
