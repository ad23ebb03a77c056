// k2cpu_build_db -- CPU ORACLE build_db front-end (test infrastructure).
// Same getopt surface as the reference build_db (src/build_db.cc:178-304).
#include "k2cpu.h"
#include <cstdlib>
#include <iostream>
#include <unistd.h>

using namespace k2cpu;

int main(int argc, char **argv) {
  BuildOptions o;
  o.spaced_seed_mask = 0;
  int opt;
  long long sig;
  while ((opt = getopt(argc, argv, "?hB:b:c:FH:m:n:o:t:k:l:M:p:r:s:S:T:XC:")) != -1) {
    switch (opt) {
      case 'h': case '?': std::cerr << "usage: k2cpu_build_db (build_db options)\n"; return 0;
      case 'B': sig = atoll(optarg); if (sig < 1) { std::cerr << "must have positive block size\n"; return 64; } o.block_size = sig; break;
      case 'b': sig = atoll(optarg); if (sig < 1) { std::cerr << "must have positive subblock size\n"; return 64; } o.subblock_size = sig; break;
      case 'r': sig = atoll(optarg); if (sig < 0 || sig > 31) { std::cerr << "bad taxid bit storage\n"; return 64; } o.requested_bits_for_taxid = sig; break;
      case 'p': o.num_threads = atoi(optarg); break;
      case 'H': o.hashtable_filename = optarg; break;
      case 'm': o.id_map_filename = optarg; break;
      case 'n': o.ncbi_taxonomy_directory = optarg; break;
      case 'o': o.options_filename = optarg; break;
      case 't': o.taxonomy_filename = optarg; break;
      case 'S': o.spaced_seed_mask = parse_bitstring(optarg); break;
      case 'T': o.toggle_mask = parse_bitstring(optarg); break;
      case 'k': sig = atoll(optarg); if (sig < 1) { std::cerr << "k must be positive integer\n"; return 64; } o.k = sig; break;
      case 'l': sig = atoll(optarg); if (sig < 1 || sig > 31) { std::cerr << "bad l\n"; return 64; } o.l = sig; break;
      case 'c': sig = atoll(optarg); if (sig < 1) { std::cerr << "capacity must be positive integer\n"; return 64; } o.capacity = sig; break;
      case 'M': sig = atoll(optarg); if (sig < 1) { std::cerr << "max capacity must be positive integer\n"; return 64; } o.maximum_capacity = sig; break;
      case 'F': o.deterministic = false; break;
      case 'X': o.input_is_protein = true; break;
      case 'C': sig = atoll(optarg); if (sig != 32) { std::cerr << "the oracle builds 32-bit cells only\n"; return 64; } break;
    }
  }
  if (o.spaced_seed_mask != 0)
    o.spaced_seed_mask = expand_spaced_seed_mask(o.spaced_seed_mask, o.input_is_protein ? 4 : 2);
  if (o.hashtable_filename.empty() || o.id_map_filename.empty() ||
      o.ncbi_taxonomy_directory.empty() || o.options_filename.empty() ||
      o.taxonomy_filename.empty()) {
    std::cerr << "missing mandatory filename parameter\n";
    return 64;
  }
  if (o.k == 0 || o.l == 0 || o.capacity == 0) { std::cerr << "missing mandatory integer parameter\n"; return 64; }
  if (o.k < o.l) { std::cerr << "k cannot be less than l\n"; return 64; }
  if (o.block_size < o.subblock_size) { std::cerr << "block size cannot be less than subblock size\n"; return 64; }
  if (o.maximum_capacity > o.capacity) { std::cerr << "maximum capacity option shouldn't specify larger capacity than normal\n"; return 64; }
  return run_build(o);
}
