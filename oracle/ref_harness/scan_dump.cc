// Harness: drives the reference MinimizerScanner (src/mmscanner.cc) and prints
// one line per NextMinimizer() return, so the oracle's scanner restatement and
// the GPU scan can be checked against the reference on the same strings.
// Input lines: k l spaced_mask toggle_mask dna(0/1) start finish sequence
// ("-" for an empty sequence; finish = -1 means end of string).
// Output: "# <n>" then n lines "<ambiguous 0/1> <minimizer hex>".
// An optional argument sets the scanner's revcom_version (default 1).
#include "mmscanner.h"
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <sstream>
#include <string>

using namespace kraken2;

int main(int argc, char **argv) {
  const int revcom_version = argc > 1 ? std::atoi(argv[1]) : 1;
  std::string line;
  while (std::getline(std::cin, line)) {
    if (line.empty()) continue;
    std::istringstream iss(line);
    long long k, l, start, finish;
    unsigned long long mask, toggle;
    int dna;
    std::string seq;
    iss >> k >> l >> std::hex >> mask >> toggle >> std::dec >> dna >> start >> finish >> seq;
    if (seq == "-") seq.clear();
    MinimizerScanner scanner(k, l, mask, dna != 0, toggle, revcom_version);
    size_t fin = finish < 0 ? SIZE_MAX : (size_t) finish;
    scanner.LoadSequence(seq, (size_t) start, fin);
    std::ostringstream out;
    size_t n = 0;
    uint64_t *m;
    while ((m = scanner.NextMinimizer()) != nullptr) {
      out << (scanner.is_ambiguous() ? 1 : 0) << " " << std::hex << *m << std::dec << "\n";
      n++;
    }
    std::cout << "# " << n << "\n" << out.str();
  }
  return 0;
}
