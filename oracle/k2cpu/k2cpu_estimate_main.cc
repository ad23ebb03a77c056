// k2cpu_estimate_capacity -- CPU ORACLE estimate_capacity front-end
// (test infrastructure).  getopt as src/estimate_capacity.cc:108-170.
#include "k2cpu.h"
#include <cstdlib>
#include <iostream>
#include <unistd.h>

using namespace k2cpu;

int main(int argc, char **argv) {
  EstimateOptions o;
  int opt;
  long long sig;
  while ((opt = getopt(argc, argv, "?hk:l:n:S:T:B:p:X")) != -1) {
    switch (opt) {
      case 'p': o.threads = atoi(optarg); break;
      case 'B': break;
      case 'n': sig = atoll(optarg); if (sig < 1 || sig > 1024) { std::cerr << "bad n\n"; return 64; } o.n = sig; break;
      case 'k': o.k = atoll(optarg); break;
      case 'l': o.l = atoll(optarg); break;
      case 'X': o.input_is_protein = true; break;
      case 'S': o.spaced_seed_mask = parse_bitstring(optarg); break;
      case 'T': o.toggle_mask = parse_bitstring(optarg); break;
      default: std::cerr << "usage: k2cpu_estimate_capacity -k K -l L [-n N] [-S mask] [-X]\n"; return 64;
    }
  }
  if (o.spaced_seed_mask != 0)
    o.spaced_seed_mask = expand_spaced_seed_mask(o.spaced_seed_mask, o.input_is_protein ? 4 : 2);
  if (o.k == 0 || o.l == 0 || o.k < o.l) { std::cerr << "bad k/l\n"; return 64; }
  std::cout << estimate_capacity(o, nullptr) << std::endl;
  return 0;
}
