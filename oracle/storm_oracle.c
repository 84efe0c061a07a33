/*
 * storm_oracle.c -- CPU restatement of the SToRM hot path (test infrastructure only).
 *
 * This file is the checker for the HIP product in storm_amd/.  It is never linked
 * into the product.  Each function cites the reference lines it restates.
 */
#include "storm_oracle.h"

#include <stdlib.h>
#include <string.h>
#include <math.h>

static void *xmalloc(size_t n) {
  void *p = malloc(n ? n : 1);
  if (!p) { fprintf(stderr, "oracle: out of memory (%zu bytes)\n", n); exit(1); }
  return p;
}
static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s ? s : 1);
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); exit(1); }
  return p;
}

/* alignment.h:24-28,51-52; main.c:37-38; index.c:8-9; read_quality.c:7 */
void orc_params_default(orc_params_t *p, int color) {
  memset(p, 0, sizeof(*p));
  p->color = color;
  p->match = 5; p->mismatch = -4; p->gap_open = -9; p->gap_extend = -4;
  p->allowed_indels = 3; p->simd_diags = 3;
  p->t = 115; p->z = 100;
  p->qual_base = 33;
  p->filter_repeats = 1; p->stdev_distance = 100;
  p->n_bytes = orc_n_bytes(3);
  p->sense_mask = ORC_SENSE_FWD | ORC_SENSE_REV;
  p->unordered_M = 1;
}

/* run.c:109-127 (SSE2 table): W = lanes of one filter alignment */
int orc_filter_lanes(int d) {
  if (d <= 1) return 2;
  if (d <= 3) return 4;
  if (d <= 7) return 8;
  return 16;
}
/* run.c:146 */
int orc_n_bytes(int d) {
  if (d <= 1) return 2;
  if (d <= 3) return 4;
  return 8;
}

/* ------------------------------------------------------------------ seeds */

/* seed.c:25-95: pattern "#-##" or "1011", optional ":p1 p2" position list */
static int seed_create(const char *pattern, orc_seed_t *s) {
  memset(s, 0, sizeof(*s));
  const char *colon = strchr(pattern, ':');
  int len = (int)(colon ? (size_t)(colon - pattern) : strlen(pattern));
  s->length = len;
  s->offsets = xmalloc(sizeof(int) * (size_t)(len > 0 ? len : 1));
  for (int i = 0; i < len; ++i) {
    char c = pattern[i];
    /* seed.h symbols: '#'/'1' match, '-'/'0' wildcard, anything else undefined */
    if (c == '#' || c == '1') s->offsets[s->weight++] = i;
    else if (c == '-' || c == '0') { /* wildcard */ }
    else { free(s->offsets); return -1; }
  }
  if (s->weight == 0) { free(s->offsets); return -1; }
  if (colon && strlen(colon) > 1) {
    const char *scan = colon + 1;
    int v;
    if (sscanf(scan, "%d", &v) == 1) {
      /* count = delimiters (' ' or ',') + 1 (seed.c:69-77) */
      int cnt = 1;
      const char *q = scan, *d;
      while (q && ((d = strchr(q, ' ')) || (d = strchr(q, ',')))) { cnt++; q = d + 1; }
      s->positions = xmalloc(sizeof(int) * (size_t)cnt);
      char *copy = strdup(scan), *save = NULL, *tok;
      int n = 0;
      for (char *ptr = copy; ; ptr = NULL) {
        tok = strtok_r(ptr, ", ", &save);
        if (!tok || n >= cnt || sscanf(tok, "%d", &s->positions[n]) != 1) break;
        n++;
      }
      s->positions_count = n;
      free(copy);
    }
  }
  return 0;
}

/* seed.c:97-135 */
int orc_seed_parse_list(const char *src, orc_seed_t **seeds) {
  *seeds = NULL;
  if (!src || !*src) return 0;
  int cap = 1;
  for (const char *q = src; (q = strchr(q, ';')); ++q) cap++;
  orc_seed_t *arr = xcalloc((size_t)cap, sizeof(orc_seed_t));
  char *copy = strdup(src), *save = NULL, *tok;
  int n = 0;
  for (char *ptr = copy; ; ptr = NULL) {
    tok = strtok_r(ptr, ";", &save);
    if (!tok) break;
    if (seed_create(tok, &arr[n]) != 0) {
      orc_seeds_free(arr, n);
      free(copy);
      return -1;
    }
    n++;
  }
  free(copy);
  *seeds = arr;
  return n;
}

void orc_seeds_free(orc_seed_t *seeds, int n) {
  if (!seeds) return;
  for (int i = 0; i < n; ++i) { free(seeds[i].offsets); free(seeds[i].positions); }
  free(seeds);
}

/* seed.c:203-214: key = codes at the '#' offsets, first one most significant */
int orc_seed_key(const orc_seed_t *seed, const uint8_t *packed, int pos) {
  int key = 0;
  for (int k = 0; k < seed->weight; ++k) key = (key << 2) | orc_nth(packed, (long)pos + seed->offsets[k]);
  return key;
}

/* ------------------------------------------------------------------ index */

/* index.c:19-195, with the contiguous (CSR) layout of index.c:157-188 */
int orc_index_build(const orc_ref_t *ref, const orc_seed_t *seed, const orc_params_t *p, orc_index_t *idx) {
  memset(idx, 0, sizeof(*idx));
  idx->index_size = 1 << (seed->weight << 1);
  idx->map_size = ref->size - seed->length + 1;
  if (idx->map_size < 0) {
    idx->map_size = 0;
    idx->first = xcalloc((size_t)idx->index_size + 1, sizeof(int));
    idx->map = xmalloc(sizeof(int));
    return 0;
  }
  int *counts = xcalloc((size_t)idx->index_size, sizeof(int));
  int *pos_key = xmalloc(sizeof(int) * (size_t)(idx->map_size + 1));
  int *pos_list = xmalloc(sizeof(int) * (size_t)(idx->map_size + 1));
  int n_pos = 0;
  long count = 0;
  if (seed->length < ref->size) {
    int r = 0;
    for (int i = 0; i < idx->map_size; ++i) {
      /* masked skip (index.c:66-72): window i overlaps (start-s .. ] of a region -> jump to end-1 */
      while (r < ref->n_regions && (i + seed->length >= ref->regions[r].start)) {
        i = ref->regions[r].end - 1;
        r++;
      }
      /* A region that reaches the end of the sequence makes the reference index position
       * end-1 >= map_size and read past its map[] (undefined); such positions are left out. */
      if (i >= idx->map_size) break;
      int key = orc_seed_key(seed, ref->seq, i);
      counts[key]++;
      pos_key[n_pos] = key;
      pos_list[n_pos] = i;
      n_pos++;
      count++;
    }
  }
  char *erased = xcalloc((size_t)idx->index_size, 1);
  if (p->filter_repeats) {
    /* index.c:119-141: same expression order, double precision */
    double avg = 0;
    for (int key = 0; key < idx->index_size; ++key)
      if (counts[key] > 0) avg += 1.0 / count * counts[key] * counts[key];
    avg /= count;
    double sd = 0;
    for (int key = 0; key < idx->index_size; ++key)
      if (counts[key] > 0) sd += (counts[key] - avg) * (counts[key] - avg) / count;
    sd = sqrt(sd);
    for (int key = 0; key < idx->index_size; ++key)
      if (counts[key] && (counts[key] - avg) > p->stdev_distance * sd) erased[key] = 1;
  }
  /* index.c:157-188: offsets of kept keys, positions ascending within a key */
  idx->first = xmalloc(sizeof(int) * ((size_t)idx->index_size + 1));
  idx->first[0] = 0;
  for (int key = 0; key < idx->index_size; ++key)
    idx->first[key + 1] = idx->first[key] + (erased[key] ? 0 : counts[key]);
  idx->map = xmalloc(sizeof(int) * (size_t)(idx->first[idx->index_size] + 1));
  int *fill = xmalloc(sizeof(int) * (size_t)idx->index_size);
  memcpy(fill, idx->first, sizeof(int) * (size_t)idx->index_size);
  for (int k = 0; k < n_pos; ++k) {
    int key = pos_key[k];
    if (!erased[key]) idx->map[fill[key]++] = pos_list[k];
  }
  free(fill); free(erased); free(counts); free(pos_key); free(pos_list);
  return 0;
}

void orc_index_free(orc_index_t *idx) {
  free(idx->first); free(idx->map);
  idx->first = idx->map = NULL;
}

/* index.c:212-241 (contiguous branch) */
int orc_index_next_hit(const orc_index_t *idx, int key, int *hit_pointer) {
  int a = *hit_pointer;
  if (a < 0) a = idx->first[key];
  else a++;
  if (a >= idx->first[key + 1]) { *hit_pointer = -1; return -1; }
  *hit_pointer = a;
  return idx->map[a];
}

/* ------------------------------------------------------------- candidates */

typedef struct { int hit_pos; int read_pos; int seed_id; } orc_heap_t;

static void heap_insert(orc_heap_t *h, int *size, int hit, int rp, int sid) {   /* run.c:481-498 */
  int u = *size, t = (u - 1) >> 1;
  while (u > 0 && hit < h[t].hit_pos) { h[u] = h[t]; u = t; t = (u - 1) >> 1; }
  h[u].hit_pos = hit; h[u].read_pos = rp; h[u].seed_id = sid;
  (*size)++;
}
static void heap_replace_top(orc_heap_t *h, int size, int hit, int rp, int sid) {   /* run.c:501-525 */
  int u = 0;
  while (u < size) {
    int umin = u, mn = hit;
    if (2 * u + 1 < size) {
      if (h[2 * u + 1].hit_pos < mn) { umin = 2 * u + 1; mn = h[2 * u + 1].hit_pos; }
      if (2 * u + 2 < size && h[2 * u + 2].hit_pos < mn) umin = 2 * u + 2;
    }
    if (umin == u) break;
    h[u] = h[umin];
    u = umin;
  }
  h[u].hit_pos = hit; h[u].read_pos = rp; h[u].seed_id = sid;
}

/* run.c:570-698: k-way heap merge of the per-(seed, read position) hit lists by diagonal,
 * keeping a diagonal only when it is above the previously popped one (which starts at -I). */
int orc_candidates(const orc_index_t *idx, const orc_seed_t *seeds, int n_seeds,
                   const uint8_t *read, int read_len, int I, int *out, int cap) {
  int total_pos = 0;
  for (int s = 0; s < n_seeds; ++s)
    total_pos += seeds[s].positions ? seeds[s].positions_count : (read_len - seeds[s].length + 1 > 0 ? read_len - seeds[s].length + 1 : 0);
  orc_heap_t *heap = xmalloc(sizeof(orc_heap_t) * (size_t)(total_pos + 1));
  int **key = xmalloc(sizeof(int *) * (size_t)n_seeds);
  int **hp = xmalloc(sizeof(int *) * (size_t)n_seeds);
  int size = 0;
  for (int s = 0; s < n_seeds; ++s) {
    int npos = seeds[s].positions ? seeds[s].positions_count : read_len - seeds[s].length + 1;
    int maxpos = read_len + 1;
    if (seeds[s].positions)
      for (int k = 0; k < npos; ++k) if (seeds[s].positions[k] + 1 > maxpos) maxpos = seeds[s].positions[k] + 1;
    key[s] = xmalloc(sizeof(int) * (size_t)maxpos);
    hp[s] = xmalloc(sizeof(int) * (size_t)maxpos);
    for (int k = 0; k < maxpos; ++k) { key[s][k] = -1; hp[s][k] = -1; }
    for (int k = 0; k < npos; ++k) {
      int rp = seeds[s].positions ? seeds[s].positions[k] : k;
      key[s][rp] = orc_seed_key(&seeds[s], read, rp);
      int q = orc_index_next_hit(&idx[s], key[s][rp], &hp[s][rp]);
      if (q >= 0) heap_insert(heap, &size, q - rp, rp, s);
    }
  }
  int n = 0, old = -I;
  while (size) {
    int hit = heap[0].hit_pos, rp = heap[0].read_pos, s = heap[0].seed_id;
    int q = orc_index_next_hit(&idx[s], key[s][rp], &hp[s][rp]);
    if (q >= 0) heap_replace_top(heap, size, q - rp, rp, s);
    else { size--; heap_replace_top(heap, size, heap[size].hit_pos, heap[size].read_pos, heap[size].seed_id); }
    if (hit > old) {
      if (n < cap) out[n] = hit;
      n++;
    }
    old = hit;
  }
  for (int s = 0; s < n_seeds; ++s) { free(key[s]); free(hp[s]); }
  free(key); free(hp); free(heap);
  return n;
}

/* ----------------------------------------------------------------- filter */

unsigned orc_filter_threshold(int z) {   /* alignment_simd.c:3643-3645 */
  unsigned u = (unsigned)z;
  if (u >= 255) u = 254;
  return u;
}

static inline int sat8(int v) { return v < 0 ? 0 : (v > 255 ? 255 : v); }

/* One candidate of alignment_sse2__align_{mono,pair,quad,octa} (alignment_simd.c:3731-4511).
 * Reference geometry: the SIMD step l = 2i + delta holds cell (row i, diagonal delta), delta in
 * [0, 2W); even steps shift in a read code, odd ones a genome code (:3782-3795).  At each cell:
 *   M0 = read==genome ? adds(Mc, match) : subs(Mc, mismatch)          (:3810-3830, Mc = corner)
 *   I  = max(subs(max(I_top, I_left), ext), subs(max(M_top, M_left), open))   (:3850-3870)
 *   M  = max(M0, I), masked to 0 outside rows [0, L) (:3879); the running max is the score.
 * top = (i-1, delta+1), left = (i, delta-1); out-of-band and row -1 neighbours are 0.
 * The row mask table (setlength, :3412-3625) equals rows [0, L) when L >= W; shorter reads
 * make the reference read read codes past the read end, so they are rejected here. */
int orc_filter_score(const uint8_t *genome, long pos, const uint8_t *read, int L, int W,
                     int match, int mismatch, int gap_open, int gap_extend) {
  if (L < W || W < 1 || W > 16) return -1;
  const int D = 2 * W;
  const int mt = abs(match) & 0xff, mm = abs(mismatch) & 0xff, op = abs(gap_open) & 0xff, ex = abs(gap_extend) & 0xff;
  int Mp[33], Ip[33], Mc[33], Ic[33];
  for (int d = 0; d <= D; ++d) { Mp[d] = Ip[d] = 0; }
  int best = 0;
  for (int i = 0; i < L; ++i) {
    int r = orc_nth(read, i);
    for (int d = 0; d < D; ++d) {
      int g = orc_nth(genome, pos + i + d);
      int corner = Mp[d];
      int m0 = (r == g) ? sat8(corner + mt) : sat8(corner - mm);
      int It = (d + 1 < D) ? Ip[d + 1] : 0, Mt = (d + 1 < D) ? Mp[d + 1] : 0;
      int Il = d ? Ic[d - 1] : 0, Ml = d ? Mc[d - 1] : 0;
      int ie = sat8((It > Il ? It : Il) - ex);
      int io = sat8((Mt > Ml ? Mt : Ml) - op);
      int Iv = ie > io ? ie : io;
      int Mv = m0 > Iv ? m0 : Iv;
      Ic[d] = Iv; Mc[d] = Mv;
      if (Mv > best) best = Mv;
    }
    memcpy(Mp, Mc, sizeof(int) * (size_t)D);
    memcpy(Ip, Ic, sizeof(int) * (size_t)D);
  }
  return best;
}

/* --------------------------------------------------------------------- DP */

typedef struct {
  int score;
  uint8_t chk;      /* color_checksum */
  char from;        /* I_MASK 0x2 | J_MASK 0x1 */
  short indels;
  short fresh;      /* color_mismatch_freshness */
  uint8_t autoc;    /* color_autocorrect */
} orc_cell_t;

#define ORC_I_MASK 0x2
#define ORC_J_MASK 0x1

/* alignment.c:255-285 (SCALE_SCORE alignment.h:30) */
static void pair_scores(const orc_params_t *p, int ps[2][4]) {
  for (int q = 0; q < 4; ++q) {
    short m = (short)p->match, x = (short)p->mismatch;
    ps[0][q] = (int)(((m) * (q + 1) + 0.) / 4);
    ps[1][q] = (int)(((x) * (q + 1) + 0.) / 4);
  }
}

/* alignment.c:465-468 */
static inline int color_adjust(uint8_t prev, uint8_t crt, uint8_t reference, int crt_score) {
  return (prev && (((prev ^ crt) == 0) || ((prev ^ crt) == reference)) && crt_score != ORC_MIN_SCORE && crt_score < 0)
         ? -2 * crt_score : crt_score;
}

/* Full matrix DP (alignment.c:326-360 init, :488-632 align).  Rows i in [0, L), columns
 * c = j - i in [0, 2I] with border cells at c = -1 and c = 2I + 1 (alignment.c:303-318). */
static int dp_fill(const uint8_t *read, const uint8_t *qlev, int L, const uint8_t *ref, const uint8_t *refmask,
                   int I, const orc_params_t *p, int min_accepted, int use_exit, int *S_out,
                   orc_cell_t *mat, int *best_j_out) {
  const int C = 2 * I + 3;   /* with two borders */
  int ps[2][4];
  pair_scores(p, ps);
  const int gap_open = (short)p->gap_open, gap_ext = (short)p->gap_extend;
#define CELL(i, c) mat[(size_t)(i) * C + (c) + 1]
  for (int i = 0; i < L; ++i) {
    CELL(i, -1).score = ORC_MIN_SCORE; CELL(i, -1).indels = (short)I; CELL(i, -1).from = 3;
    CELL(i, 2 * I + 1).score = ORC_MIN_SCORE; CELL(i, 2 * I + 1).indels = (short)I; CELL(i, 2 * I + 1).from = 3;
    CELL(i, -1).chk = CELL(i, 2 * I + 1).chk = 0; CELL(i, -1).fresh = CELL(i, 2 * I + 1).fresh = 0;
    CELL(i, -1).autoc = CELL(i, 2 * I + 1).autoc = 0;
    for (int c = 0; c <= 2 * I; ++c) {
      orc_cell_t *x = &CELL(i, c);
      int j = i + c;
      if (refmask[j] == 0) {
        if (p->color) {
          x->autoc = x->chk = (uint8_t)(read[i] ^ ref[j]);
          x->fresh = x->autoc ? qlev[i] : 0;
          x->score = ps[x->chk ? 1 : 0][qlev[i]];
        } else {
          x->chk = 0; x->autoc = 0; x->fresh = 0;
          x->score = ps[read[i] != ref[j] ? 1 : 0][qlev[i]];
        }
      } else {
        x->chk = x->autoc = 0; x->fresh = 0;
        x->score = ORC_MIN_SCORE;
      }
      x->from = ORC_I_MASK | ORC_J_MASK;
      x->indels = 0;
    }
  }
  const int top_match = ps[0][3];
  int best_possible = (p->color ? (L - 2) : L) * top_match;
  int S = 0x7fffffff;
  for (int i = 1; i < L; ++i) {
    int max_on_line = ORC_MIN_SCORE;
    for (int c = 0; c <= 2 * I; ++c) {
      orc_cell_t *crt = &CELL(i, c);
      orc_cell_t *corner = &CELL(i - 1, c);
      orc_cell_t *top = &CELL(i - 1, c + 1);
      orc_cell_t *left = &CELL(i, c - 1);
      orc_cell_t *it = corner;
      int j = i + c;
      int match_score = p->color ? color_adjust(corner->chk, crt->chk, corner->autoc, crt->score) : crt->score;
      crt->score = corner->score + match_score;
      int tmp;
      if (top->indels < I && (tmp = top->score + ((top->from ^ ORC_I_MASK) ? gap_open : gap_ext)) > crt->score) {
        it = top; crt->score = tmp;
      }
      if (left->indels < I && (tmp = left->score + ((left->from ^ ORC_J_MASK) ? gap_open : gap_ext)) > crt->score) {
        it = left; crt->score = tmp;
      }
      if (it == corner) {
        crt->indels = it->indels;
        if (p->color) {   /* alignment.c:533-578 */
          crt->chk = crt->chk ^ it->chk;
          if (it->chk != crt->chk) {
            if (it->chk == 0) {
              crt->autoc = crt->chk;
              crt->fresh = qlev[i];
            } else if (crt->chk == it->autoc) {
              crt->autoc = 0; crt->chk = 0; crt->fresh = 0;
            } else if (crt->chk == 0) {
              if (it->fresh - 1 > 0) { crt->autoc = it->autoc; crt->fresh = (short)(it->fresh - 1); }
              else { crt->autoc = 0; crt->fresh = 0; }
            } else {
              if (it->fresh - 1 > 0) { crt->autoc = it->autoc; crt->fresh = (short)(it->fresh - 1); }
              else { crt->autoc = crt->chk; crt->fresh = qlev[i]; }
            }
          } else {
            if (it->autoc != 0) {
              if (it->fresh - 1 > 0) {
                if (crt->chk == 0) { crt->autoc = 0; crt->fresh = 0; }
                else { crt->autoc = it->autoc; crt->fresh = (short)(it->fresh - 1); }
              }
            } else {
              crt->chk = 0; crt->autoc = 0; crt->fresh = 0;
            }
          }
        }
      } else if (it == top) {   /* alignment.c:580-589 */
        crt->from = ORC_I_MASK;
        crt->indels = (short)(it->indels + 1);
        if (p->color) {
          crt->chk = (uint8_t)(top->chk ^ read[i]);
          if (it->fresh - 1 > 0) { crt->fresh = (short)(it->fresh - 1); crt->autoc = corner->autoc; }
        }
      } else {   /* alignment.c:590-599 */
        crt->from = ORC_J_MASK;
        crt->indels = (short)(it->indels + 1);
        if (p->color) {
          crt->chk = (uint8_t)(left->chk ^ ref[j]);
          if (it->fresh - 1 > 0) { crt->fresh = (short)(it->fresh - 1); crt->autoc = corner->autoc; }
        }
      }
      if (max_on_line < crt->score) max_on_line = crt->score;
    }
    if (max_on_line + best_possible < S) S = max_on_line + best_possible;
    if (use_exit && max_on_line + best_possible < min_accepted) { if (S_out) *S_out = S; return ORC_MIN_SCORE; }   /* :613-616 */
    best_possible -= top_match;
  }
  int best = ORC_MIN_SCORE, bj = -1;
  for (int c = 0; c <= 2 * I; ++c) {   /* :621-630, leftmost maximum of the last row */
    if (best < CELL(L - 1, c).score) { best = CELL(L - 1, c).score; bj = L - 1 + c; }
  }
  if (S_out) *S_out = S;
  *best_j_out = bj;
  return best;
#undef CELL
}

int orc_align(const uint8_t *read, const uint8_t *qlev, int L, const uint8_t *ref, const uint8_t *refmask,
              int I, const orc_params_t *p, int min_accepted, orc_align_out_t *out) {
  const int C = 2 * I + 3;
  orc_cell_t *mat = xmalloc(sizeof(orc_cell_t) * (size_t)L * (size_t)C);
  int bj = -1;
  int score = dp_fill(read, qlev, L, ref, refmask, I, p, min_accepted, 1, NULL, mat, &bj);
  out->score = score;
  out->best_i = out->best_j = out->best_i0 = out->best_j0 = -1;
  out->tb_len = 0;
  if (score != ORC_MIN_SCORE || bj >= 0) {
    if (bj >= 0) { out->best_i = L - 1; out->best_j = bj; }
  }
  if (score > min_accepted && bj >= 0) {
    /* alignment.c:642-734: walk `from` back to row 0, emitting 4-bit codes */
#define CELLP(i, c) (&mat[(size_t)(i) * C + (c) + 1])
    int i = out->best_i, j = out->best_j;
    uint8_t color = 0;
    int n = 0;
    while (i >= 0 && i < L && j >= i && j <= i + 2 * I) {
      orc_cell_t *cell = CELLP(i, j - i);
      out->best_i0 = i; out->best_j0 = j;
      int di = -((cell->from & ORC_I_MASK) >> 1);
      int dj = -(cell->from & ORC_J_MASK);
      uint8_t code;
      if (di && !dj) code = (uint8_t)(0xC | read[i]);
      else if (dj && !di) code = 0x1;
      else {
        if (read[i] != ref[j]) {
          if (!p->color) code = (uint8_t)(0x4 | read[i]);
          else if (color || (i != 0 && j != 0 && CELLP(i, j - i)->score > CELLP(i - 1, j - i)->score)) {
            code = (uint8_t)(0x8 | read[i]);
            color = (uint8_t)(color ^ (read[i] ^ ref[j]));
          } else code = (uint8_t)(0x4 | read[i]);
        } else code = 0x3;
      }
      if (n < (int)sizeof(out->tb)) out->tb[n] = code;
      n++;
      i += di; j += dj;
    }
    out->tb_len = n;
#undef CELLP
  }
  free(mat);
  return score;
}

void orc_align_bound(const uint8_t *read, const uint8_t *qlev, int L, const uint8_t *ref, const uint8_t *refmask,
                     int I, const orc_params_t *p, int *score, int *S) {
  const int C = 2 * I + 3;
  orc_cell_t *mat = xmalloc(sizeof(orc_cell_t) * (size_t)L * (size_t)C);
  int bj;
  *score = dp_fill(read, qlev, L, ref, refmask, I, p, 0, 0, S, mat, &bj);
  free(mat);
}

/* ---------------------------------------------------------------- hit map */

/* hit_map.c:50-144.  `row` holds MAP_DETAIL slots; the traceback of `al` is stored
 * reversed (as the DP wrote it) and is copied forward. */
int orc_hitmap_update(orc_hit_t *row, int ref_id, int ref_pos, const orc_align_out_t *al, int score, int sense) {
  int k;
  for (k = 0; k < ORC_MAP_DETAIL && row[k].ref_id != ORC_UNMAPPED; ++k) {
    if (row[k].ref_id == ref_id &&
        ((row[k].ref_start == ref_pos + al->best_j0 && row[k].read_start == al->best_i0) ||
         (row[k].ref_end == ref_pos + al->best_j && row[k].read_end == al->best_i))) {
      if (score > row[k].score) break;
      return ORC_UNMAPPED;
    }
  }
  if (k == ORC_MAP_DETAIL) k--;
  int i = k;
  while (i >= 0 && score > row[i].score) i--;
  if (i == ORC_MAP_DETAIL - 1) return ORC_UNMAPPED;
  free(row[k].traceback);
  row[k].traceback = NULL;
  for (int j = k; j > i + 1; --j) row[j] = row[j - 1];
  orc_hit_t *h = &row[i + 1];
  h->score = score;
  h->ref_id = ref_id;
  h->ref_start = ref_pos + al->best_j0;
  h->ref_end = ref_pos + al->best_j;
  h->read_start = al->best_i0;
  h->read_end = al->best_i;
  h->sense = (sense & ORC_SENSE_REV) ? -1 : 1;
  h->traceback_len = al->tb_len;
  h->traceback = xcalloc((size_t)(al->tb_len + 1) / 2 + 1, 1);
  for (int q = 0; q < al->tb_len; ++q) {
    uint8_t code = al->tb[al->tb_len - 1 - q] & 0xF;
    /* TRACEBACK_SEQ_SET (alignment.h:83): even index -> high nibble */
    if (q & 1) h->traceback[q >> 1] = (uint8_t)((h->traceback[q >> 1] & 0xF0) | code);
    else h->traceback[q >> 1] = (uint8_t)((h->traceback[q >> 1] & 0x0F) | (code << 4));
  }
  return i + 1;
}

static orc_hit_t *hitmap_row_new(void) {
  orc_hit_t *row = xcalloc(ORC_MAP_DETAIL, sizeof(orc_hit_t));
  for (int k = 0; k < ORC_MAP_DETAIL; ++k) {
    row[k].score = ORC_MIN_SCORE;
    row[k].ref_id = row[k].ref_start = row[k].ref_end = ORC_UNMAPPED;
    row[k].read_start = row[k].read_end = ORC_UNMAPPED;
    row[k].sense = 0;
  }
  return row;
}

/* ------------------------------------------------------------ process_read */

/* run.c:555-747 for one read, one strand, one reference sequence */
static void process_read(const orc_read_t *read, int L, long read_id, int ref_id, const orc_ref_t *ref,
                         const orc_index_t *idx, const orc_seed_t *seeds, int n_seeds,
                         const orc_params_t *p, int sense, orc_result_t *res,
                         int *cand_buf, int cand_cap, uint8_t *rcodes, uint8_t *qlev,
                         uint8_t *wref, uint8_t *wmask) {
  const int I = p->allowed_indels, d = p->simd_diags, W = orc_filter_lanes(d);
  const unsigned thr = orc_filter_threshold(p->z);
  int T = p->t;
  int n = orc_candidates(idx, seeds, n_seeds, read->seq, L, I, cand_buf, cand_cap);
  if (n > cand_cap) { fprintf(stderr, "oracle: candidate buffer too small\n"); exit(1); }
  res->n_candidates += n;
  for (int c = 0; c < n; ++c) {
    int h = cand_buf[c];
    int fs = orc_filter_score(ref->seq, (long)h - d, read->seq, L, W, p->match, p->mismatch, p->gap_open, p->gap_extend);
    if (!(fs >= 0 && (unsigned)fs > thr)) continue;
    res->n_filter_pass++;
    int win = h - I > -d ? h - I : -d;   /* run.c:667 */
    for (int k = 0; k < L + 2 * I; ++k) {
      wref[k] = (uint8_t)orc_nth(ref->seq, (long)win + k);
      wmask[k] = (uint8_t)orc_nth(ref->mask, (long)win + k);
    }
    orc_align_out_t al;
    res->n_dp++;
    int score = orc_align(rcodes, qlev, L, wref, wmask, I, p, T, &al);
    if (score > T) {
      res->n_accepted++;
      if (!res->map[read_id]) { res->map[read_id] = hitmap_row_new(); res->mapped++; }
      orc_hitmap_update(res->map[read_id], ref_id, win, &al, score, sense);
      if (res->map[read_id][ORC_MAP_DETAIL - 1].score > T) T = res->map[read_id][ORC_MAP_DETAIL - 1].score;
    }
  }
}

/* run.c:764-1017 without the ordered genome map: -M mode */
int orc_map(const orc_reads_t *db, const orc_ref_t *refs, int n_refs,
            const orc_seed_t *seeds, int n_seeds, const orc_params_t *p, orc_result_t *res) {
  memset(res, 0, sizeof(*res));
  res->n_reads = db->n;
  res->map = xcalloc((size_t)(db->n ? db->n : 1), sizeof(orc_hit_t *));
  const int L = db->read_len, I = p->allowed_indels;
  orc_index_t *idx = xcalloc((size_t)n_seeds, sizeof(orc_index_t));
  int cand_cap = 1 << 22;
  int *cand = xmalloc(sizeof(int) * (size_t)cand_cap);
  uint8_t *rcodes = xmalloc((size_t)L + 1), *qlev = xmalloc((size_t)L + 1);
  uint8_t *wref = xmalloc((size_t)(L + 2 * I + 1)), *wmask = xmalloc((size_t)(L + 2 * I + 1));
  size_t rbytes = (size_t)(L / 4 + 1 + 64);
  orc_read_t rev;
  memset(&rev, 0, sizeof(rev));
  rev.seq = xcalloc(rbytes, 1);
  rev.qual = xcalloc(rbytes, 1);
  for (int r = 0; r < n_refs; ++r) {
    for (int s = 0; s < n_seeds; ++s) orc_index_build(&refs[r], &seeds[s], p, &idx[s]);
    for (int pass = 0; pass < 2; ++pass) {
      int sense = pass == 0 ? ORC_SENSE_FWD : ORC_SENSE_REV;
      if (!(p->sense_mask & sense)) continue;
      for (long id = 0; id < db->n; ++id) {
        const orc_read_t *rd = &db->reads[id];
        const orc_read_t *use = rd;
        if (sense == ORC_SENSE_REV) {
          memset(rev.seq, 0, rbytes);
          memset(rev.qual, 0, rbytes);
          orc_read_reverse(rd, &rev, L, p->color);
          use = &rev;
        }
        for (int k = 0; k < L; ++k) {
          rcodes[k] = (uint8_t)orc_nth(use->seq, k);
          qlev[k] = rd->qual ? (uint8_t)orc_nth(use->qual, k) : 3;   /* alignment.c:151-162 */
        }
        process_read(use, L, id, r, &refs[r], idx, seeds, n_seeds, p, sense, res, cand, cand_cap,
                     rcodes, qlev, wref, wmask);
      }
    }
    for (int s = 0; s < n_seeds; ++s) orc_index_free(&idx[s]);
  }
  free(rev.seq); free(rev.qual);
  free(cand); free(rcodes); free(qlev); free(wref); free(wmask); free(idx);
  return 0;
}

void orc_result_free(orc_result_t *res) {
  if (!res->map) return;
  for (long i = 0; i < res->n_reads; ++i) {
    if (!res->map[i]) continue;
    for (int k = 0; k < ORC_MAP_DETAIL; ++k) free(res->map[i][k].traceback);
    free(res->map[i]);
  }
  free(res->map);
  res->map = NULL;
}
