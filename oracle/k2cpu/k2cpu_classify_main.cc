// k2cpu_classify -- CPU ORACLE classify front-end (test infrastructure).
// Same getopt surface as the reference classify (src/classify.cc:1073-1168).
#include "k2cpu.h"
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <unistd.h>

using namespace k2cpu;

static void usage(int code) {
  std::cerr << "Usage: k2cpu_classify [options] <fasta/fastq file(s)>\n"
               "  -H hash.k2d -t taxo.k2d -o opts.k2d (mandatory)\n"
               "  -q -T NUM -p NUM -Q NUM -P -S -R file -z -n -g NUM -C file -U file -O file -K\n";
  exit(code);
}

int main(int argc, char **argv) {
  ClassifyOptions o;
  int opt;
  while ((opt = getopt(argc, argv, "h?H:t:o:T:p:R:C:U:O:Q:g:nmzqPSMKD")) != -1) {
    switch (opt) {
      case 'h': case '?': usage(0); break;
      case 'H': o.index_filename = optarg; break;
      case 't': o.taxonomy_filename = optarg; break;
      case 'T':
        o.confidence_threshold = std::stod(optarg);
        if (o.confidence_threshold < 0 || o.confidence_threshold > 1) {
          std::cerr << "k2cpu_classify: confidence threshold must be in [0, 1]\n";
          return 64;
        }
        break;
      case 'o': o.options_filename = optarg; break;
      case 'q': o.quick_mode = true; break;
      case 'p':
        o.num_threads = atoi(optarg);
        if (o.num_threads < 1) {
          std::cerr << "k2cpu_classify: number of threads can't be less than 1\n";
          return 64;
        }
        break;
      case 'g': o.minimum_hit_groups = atoi(optarg); break;
      case 'P': o.paired_end_processing = true; break;
      case 'S': o.paired_end_processing = true; o.single_file_pairs = true; break;
      case 'm': o.mpa_style_report = true; break;
      case 'K': o.report_kmer_data = true; break;
      case 'R': o.report_filename = optarg; break;
      case 'z': o.report_zero_counts = true; break;
      case 'C': o.classified_output_filename = optarg; break;
      case 'U': o.unclassified_output_filename = optarg; break;
      case 'O': o.kraken_output_filename = optarg; break;
      case 'n': o.print_scientific_name = true; break;
      case 'Q': o.minimum_quality_score = atoi(optarg); break;
      case 'M': o.use_memory_mapping = true; break;
      case 'D': std::cerr << "k2cpu_classify: daemon mode is not part of the oracle\n"; return 64;
    }
  }
  if (o.index_filename.empty() || o.taxonomy_filename.empty() || o.options_filename.empty()) {
    std::cerr << "k2cpu_classify: mandatory filename missing\n";
    usage(64);
  }
  if (o.mpa_style_report) {
    std::cerr << "k2cpu_classify: mpa-style reports are not part of the oracle\n";
    return 64;
  }
  for (int i = optind; i < argc; i++) o.filenames.push_back(argv[i]);
  std::cerr << "Loading database information...";
  Index idx;
  std::string err;
  if (!load_index(o, idx, err)) {
    std::cerr << "\nk2cpu_classify: " << err << "\n";
    return 65;
  }
  o.use_translated_search = !idx.opts.dna_db;
  std::cerr << " done." << std::endl;
  ClassifyStats stats;
  TaxonCounters counters;
  double secs = 0;
  int rc = run_classify(o, idx, stats, counters, &secs);
  if (rc) return rc;
  report_stats(secs, stats);
  if (!o.report_filename.empty())
    report_kraken_style(o.report_filename, o.report_zero_counts, o.report_kmer_data, idx.tax,
                        counters, stats.total_sequences,
                        stats.total_sequences - stats.total_classified);
  return 0;
}
