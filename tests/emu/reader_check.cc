// Prints every record of a file twice over: once through Reader::next and
// once through Reader::next_records with the given chunk size and threads
// (each chunk's records are handed back as the pool the next chunk parses
// into), so a test can compare the two record streams.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "k2g_host.h"

static void put(const k2host::Record &r) {
  printf("%d|%s|%s|%s|%s\n", r.fastq ? 1 : 0, r.header.c_str(), r.comment.c_str(), r.seq.c_str(),
         r.quals.c_str());
}

int main(int argc, char **argv) {
  if (argc != 4) return 64;
  const size_t chunk = strtoull(argv[2], nullptr, 10);
  const int threads = atoi(argv[3]);
  {
    k2host::Reader r;
    if (!r.open(argv[1])) return 66;
    k2host::Record rec;
    while (r.next(rec)) put(rec);
  }
  printf("----\n");
  k2host::Reader r;
  if (!r.open(argv[1])) return 66;
  std::vector<k2host::Record> out, pool;  // printed records go back to the pool for reuse
  while (r.next_records(chunk, threads, out, &pool)) {
    for (auto &rec : out) {
      put(rec);
      pool.push_back(std::move(rec));
    }
    out.clear();
  }
  return 0;
}
