// Runs many WorkPool::run calls back to back, each with its own function and
// item count, and checks that every item of every run was done exactly once
// by that run's function (a worker waking late must not carry an old run's
// function into a new run).
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "k2g_host.h"

int main(int argc, char **argv) {
  const int threads = argc > 1 ? atoi(argv[1]) : 8;
  const int runs = argc > 2 ? atoi(argv[2]) : 20000;
  k2host::WorkPool pool(threads);
  std::vector<std::atomic<int>> hits(4096);
  for (int r = 0; r < runs; r++) {
    const size_t n = 1 + (size_t) (r * 7919u) % 4096;
    for (size_t i = 0; i < n; i++) hits[i].store(0);
    const int tag = r;
    std::atomic<int> wrong{0};
    pool.run(n, [&, tag](size_t i) {
      if (tag != r) wrong++;
      hits[i].fetch_add(1);
    });
    for (size_t i = 0; i < n; i++)
      if (hits[i].load() != 1 || wrong.load()) {
        printf("run %d item %zu done %d times\n", r, i, hits[i].load());
        return 1;
      }
  }
  printf("ok %d runs\n", runs);
  return 0;
}
