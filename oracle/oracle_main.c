/*
 * oracle_main.c -- command line of the CPU restatement (test infrastructure only).
 * Takes the reference's flags (main.c:188) and writes the same SAM: the -M hit map output, or
 * without -M the ordered (genome map) output, contextual or greedy (-G).
 * Built twice: -DORC_COLOR=0 (nucleotide) and -DORC_COLOR=1 (color space).
 */
#include "storm_oracle.h"

#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#ifndef ORC_COLOR
#define ORC_COLOR 0
#endif

int main(int argc, char **argv) {
  orc_params_t p;
  orc_params_default(&p, ORC_COLOR);
  const char *genome = NULL, *reads = NULL, *quals = NULL, *outpath = NULL, *unmapped = NULL;
  const char *seeds = ORC_COLOR ? "####----##--##-####;####-###--#----####;####-####-####"
                                : "####--#-#--##--####;#####--###-#-###;###-##-#######";   /* seed.h:29-32 */
  int opt;
  p.unordered_M = 0;
  while ((opt = getopt(argc, argv, "g:r:q:s:b:m:x:d:e:t:z:i:v:o:O:l:u:M:N:jFRAG")) != -1) {
    switch (opt) {
      case 'g': genome = optarg; break;
      case 'r': reads = optarg; break;
      case 'q': quals = optarg; break;
      case 's': seeds = optarg; break;
      case 'b': p.qual_base = atoi(optarg); break;
      case 'm': p.match = atoi(optarg); break;
      case 'x': p.mismatch = atoi(optarg); break;
      case 'd': p.gap_open = atoi(optarg); break;
      case 'e': p.gap_extend = atoi(optarg); break;
      case 't': p.t = atoi(optarg); break;
      case 'z': p.z = atoi(optarg); break;
      case 'i':
        p.allowed_indels = atoi(optarg);
        p.simd_diags = p.allowed_indels >= 16 ? 15 : p.allowed_indels;   /* main.c:302-313 */
        break;
      case 'o': outpath = optarg; break;
      case 'O': unmapped = (!strcmp(optarg, "NULL") || !strcmp(optarg, "null")) ? NULL : optarg; break;   /* main.c:345-355 */
      case 'l':
        p.filter_repeats = 1; p.stdev_distance = atoi(optarg);
        if (p.stdev_distance <= 0) { p.filter_repeats = 0; p.stdev_distance = 0; }
        break;
      case 'u': p.uniqueness = 1; p.uniqueness_threshold = atoi(optarg); break;
      case 'M': p.unordered_M = atoi(optarg); break;
      case 'j': p.ignore_lowercase = 1; break;
      case 'F': p.sense_mask = ORC_SENSE_FWD; break;
      case 'R': p.sense_mask = ORC_SENSE_REV; break;
      case 'A':
        if (!p.unordered_M) { fprintf(stderr, "-A needs the unordered mapping (-M <number>) first\n"); return 1; }   /* main.c:377-382 */
        p.print_all = 1;
        break;
      case 'G': p.greedy = 1; break;
      case 'v': case 'N': break;
      default: return 1;
    }
  }
  if (!genome || !reads || p.unordered_M < 0 || p.unordered_M > 4) {
    fprintf(stderr, "usage: %s -g genome.fa -r reads.fq [-M <1..4>] [reference flags]\n", argv[0]);
    return 1;
  }
  p.n_bytes = orc_n_bytes(p.simd_diags);
  orc_seed_t *sd = NULL;
  int ns = orc_seed_parse_list(seeds, &sd);
  if (ns <= 0) { fprintf(stderr, "invalid seeds\n"); return -1; }
  orc_reads_t db;
  if (orc_load_reads(reads, quals, &p, &db) <= 0) { fprintf(stderr, "reads could not be loaded\n"); return -1; }
  orc_sort_reads(&db);
  orc_ref_t *refs = NULL;
  int nr = orc_load_reference(genome, db.read_len, &p, &refs);
  if (nr <= 0) { fprintf(stderr, "reference could not be loaded\n"); return -1; }
  orc_result_t res;
  orc_map(&db, refs, nr, sd, ns, &p, &res);
  FILE *out = outpath ? fopen(outpath, "w") : stdout;
  if (!out) { fprintf(stderr, "cannot open %s\n", outpath); return 1; }
  if (p.unordered_M) orc_write_sam(out, &res, &db, refs, nr, &p, argc, argv);
  else orc_write_sam_ordered(out, &res, &db, refs, nr, &p, argc, argv);
  if (out != stdout) fclose(out);
  if (unmapped) {   /* run.c:1078-1086 */
    FILE *uf = fopen(unmapped, "w");
    if (!uf) { fprintf(stderr, "cannot open %s\n", unmapped); return 1; }
    orc_write_unmapped_fastq(uf, &res, &db, &p);
    fclose(uf);
  }
  fprintf(stderr, "oracle: %ld of %ld reads mapped; candidates %ld, filter passes %ld, accepted %ld\n",
          res.mapped, res.n_reads, res.n_candidates, res.n_filter_pass, res.n_accepted);
  orc_result_free(&res);
  orc_refs_free(refs, nr);
  orc_reads_free(&db);
  orc_seeds_free(sd, ns);
  return 0;
}
