// k2g_estimate_capacity: the reference's estimate_capacity command line
// (getopt string of src/estimate_capacity.cc:108; defaults :52-60) with the
// distinct-minimizer count computed on the GPU (k2g_estimate_capacity).
// Reads sequences from stdin and prints the estimate.
#include <getopt.h>
#include <sysexits.h>

#include <cstdlib>
#include <iostream>

#include "k2g.h"
#include "k2g_host.h"

namespace {

[[noreturn]] void usage(int code = EX_USAGE) {
  std::cerr << "Usage: k2g_estimate_capacity <options>\n"
               "\n"
               "Options (*mandatory):\n"
               "* -k INT        Set length of k-mers\n"
               "* -l INT        Set length of minimizers\n"
               "  -n INT        Set maximum qualifying hash code\n"
               "  -X            Input sequences are proteins\n"
               "  -S BITSTRING  Spaced seed mask\n"
               "  -T BITSTRING  Minimizer ordering toggle mask\n"
               "  -B INT        Read block size (accepted; whole input is scanned)\n"
               "  -p INT        Number of threads (accepted; the GPU scans)" << std::endl;
  std::exit(code);
}

[[noreturn]] void die(int code, const std::string &msg) {
  std::cerr << "k2g_estimate_capacity: " << msg << std::endl;
  std::exit(code);
}

}  // namespace

int main(int argc, char **argv) {
  uint64_t k = 0, l = 0, n = 4, mask = 0, toggle = 0xe37e28c4271b5a2dULL;
  bool protein = false;
  int opt;
  while ((opt = getopt(argc, argv, "?hk:l:n:S:T:B:p:X")) != -1) {
    long long v;
    switch (opt) {
      case 'h':
      case '?': usage(0);
      case 'p':
        if (std::atoll(optarg) < 1) die(EX_USAGE, "must have at least 1 thread");
        break;
      case 'B':
        if (std::atoll(optarg) < 1) die(EX_USAGE, "block size must be positive");
        break;
      case 'n':
        v = std::atoll(optarg);
        if (v < 1) die(EX_USAGE, "n must be positive integer");
        if (v > 1024) die(EX_USAGE, "n must be no more than 1024");
        n = (uint64_t) v;
        break;
      case 'k':
        if ((v = std::atoll(optarg)) < 1) die(EX_USAGE, "k must be positive integer");
        k = (uint64_t) v;
        break;
      case 'l':
        v = std::atoll(optarg);
        if (v < 1) die(EX_USAGE, "l must be positive integer");
        if (v > 31) die(EX_USAGE, "l must be no more than 31");
        l = (uint64_t) v;
        break;
      case 'X': protein = true; break;
      case 'S': mask = std::strtoull(optarg, nullptr, 2); break;
      case 'T': toggle = std::strtoull(optarg, nullptr, 2); break;
    }
  }
  if (mask != 0) mask = k2host::expand_spaced_seed_mask(mask, protein ? 4 : 2);
  if (k == 0 || l == 0) {
    std::cerr << "missing mandatory integer parameter" << std::endl;
    usage();
  }
  if (k < l) {
    std::cerr << "k cannot be less than l" << std::endl;
    usage();
  }

  k2host::Reader reader;
  if (!reader.open(nullptr)) die(EX_NOINPUT, "unable to read sequences from standard input");
  std::string seq;
  std::vector<uint64_t> off{0};
  k2host::Record rec;
  while (reader.next(rec)) {
    // src/estimate_capacity.cc:199-201: protein sequences end in '*'
    if (protein && (rec.seq.empty() || rec.seq.back() != '*')) rec.seq.push_back('*');
    seq += rec.seq;
    off.push_back(seq.size());
  }
  k2g_build_params bp{};
  bp.k = k;
  bp.l = l;
  bp.spaced_seed_mask = mask;
  bp.toggle_mask = toggle;
  bp.dna_db = protein ? 0 : 1;
  bp.revcom_version = 1;
  uint64_t est = 0;
  const int rc = k2g_estimate_capacity(&bp, seq.data(), off.data(), off.size() - 1, (uint32_t) n, &est);
  if (rc != K2G_OK) die(rc, k2g_last_error());
  std::cout << est << std::endl;
  return 0;
}
