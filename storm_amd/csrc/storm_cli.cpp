// storm_cli.cpp -- command line with the reference's flags (main.c:188-433), running the gfx950
// pipeline through the C ABI: -M <1..4> writes the hit map SAM, no -M the ordered (genome map)
// SAM, contextual or greedy (-G).  Built as storm-gpu-nucleotide / storm-gpu-color.
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>

#include "../../include/storm_gpu.h"

#ifndef STORM_CLI_COLOR
#define STORM_CLI_COLOR 0
#endif

// seed__string_from_file (seed.c:140-179): one or more seeds per line, lines starting with '%'
// ignored, trailing blanks removed, a ';' appended to a line that does not end with one
static bool seeds_from_file(const char* path, std::string& out) {
  std::ifstream f(path);
  if (!f) return false;
  out.clear();
  std::string line;
  while (std::getline(f, line)) {
    if (!line.empty() && line[0] == '%') continue;
    while (!line.empty() && (line.back() == '\r' || line.back() == ' ' || line.back() == '\t')) line.pop_back();
    out += line;
    if (!out.empty() && out.back() != ';') out += ';';
  }
  return true;
}

int main(int argc, char** argv) {
  storm_params p;
  storm_params_default(&p, STORM_CLI_COLOR);
  const char *genome = nullptr, *reads = nullptr, *quals = nullptr, *out = "/dev/stdout";
  int mode = STORM_MODE_BATCHED, device = 0, opt;
  std::string seed_file_text;
  p.max_hits = 0;
  while ((opt = getopt(argc, argv, "g:r:q:s:S:b:m:x:d:e:t:z:i:v:o:O:l:u:M:N:jFRAGPD:")) != -1) {
    switch (opt) {
      case 'g': genome = optarg; break;
      case 'r': reads = optarg; break;
      case 'q': quals = optarg; break;
      case 's': p.seeds = optarg; break;
      case 'S':   // main.c:252-266 (the reference then asks whether to go on with the default seeds)
        if (!seeds_from_file(optarg, seed_file_text)) {
          fprintf(stderr, "%s: unable to read seeds input file %s\n", argv[0], optarg);
          return 1;
        }
        p.seeds = seed_file_text.c_str();
        break;
      case 'O':   // main.c:345-355
        p.unmapped_path = (!strcmp(optarg, "NULL") || !strcmp(optarg, "null")) ? nullptr : optarg;
        break;
      case 'b': p.qual_base = atoi(optarg); break;
      case 'm': p.match = atoi(optarg); break;
      case 'x': p.mismatch = atoi(optarg); break;
      case 'd': p.gap_open = atoi(optarg); break;
      case 'e': p.gap_extend = atoi(optarg); break;
      case 't': p.t = atoi(optarg); break;
      case 'z': p.z = atoi(optarg); break;
      case 'i': p.allowed_indels = atoi(optarg); break;
      case 'o': out = optarg; break;
      case 'l': p.stdev_distance = atoi(optarg); break;
      case 'u': p.uniqueness = atoi(optarg); break;
      case 'M': p.max_hits = atoi(optarg); break;
      case 'j': p.ignore_lowercase = 1; break;
      case 'F': p.sense_mask = 1; break;
      case 'R': p.sense_mask = 2; break;
      case 'A':
        if (!p.max_hits) {   // main.c:377-382
          fprintf(stderr, "%s: -A (output all reads) needs the unordered mapping (-M <number>) first\n", argv[0]);
          return 1;
        }
        p.print_all = 1;
        break;
      case 'G': p.greedy = 1; break;
      case 'P': mode = STORM_MODE_PLAIN; break;     // plain per-batch path
      case 'D': device = atoi(optarg); break;
      case 'v': case 'N': break;
      default: return 1;
    }
  }
  if (!genome || !reads || p.max_hits < 0 || p.max_hits > 4) {
    fprintf(stderr, "usage: %s -g genome.fa -r reads.{fq,fa,csfasta} [-q reads.qual] [-M <1..4>] [-G] [reference flags] [-P]\n", argv[0]);
    return 1;
  }
  storm_map_stats st{};
  int rc = storm_map_files(device, genome, reads, quals, &p, mode, out, argc, argv, &st);
  if (rc) {
    fprintf(stderr, "%s: %s\n", argv[0], storm_last_error());
    return 2;
  }
  fprintf(stderr, "mapped %lld of %lld reads; candidates %lld, filter passes %lld; load %.3fs index %.3fs map %.3fs sam %.3fs\n",
          (long long)st.n_mapped, (long long)st.n_reads, (long long)st.n_candidates, (long long)st.n_filter_pass,
          st.seconds_load, st.seconds_index, st.seconds_map, st.seconds_sam);
  return 0;
}
