// Harness: drives the reference TranslateToAllFrames (src/aa_translate.cc)
// and prints the six frames of each input DNA line, one per line.
#include "aa_translate.h"
#include <iostream>
#include <string>
#include <vector>

using namespace kraken2;

int main() {
  std::string line;
  std::vector<std::string> frames(6);
  while (std::getline(std::cin, line)) {
    if (line == "-") line.clear();
    TranslateToAllFrames(line, frames);
    for (int i = 0; i < 6; i++) std::cout << (frames[i].empty() ? "-" : frames[i]) << "\n";
  }
  return 0;
}
