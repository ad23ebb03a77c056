// Prints every record of a file four times over: through Reader::next,
// through Reader::next_records with the given chunk size and threads (each
// chunk's records are handed back as the pool the next chunk parses into),
// and through Reader::next_flat on a WorkPool (the flat arrays kept over two
// chunks, then cleared, so records are appended at nonzero offsets too) and
// through Reader::next_inplace, so a test can compare the record streams.
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
  printf("----\n");
  k2host::Reader rf;
  if (!rf.open(argv[1])) return 66;
  k2host::WorkPool workers(threads);  // the flat parse runs its parts on a pool
  rf.set_pool(&workers);
  std::vector<k2host::FlatRec> recs;
  std::vector<char> text, bases;
  k2host::SeqSink sink;
  sink.grow = [&](k2host::SeqSink &s, size_t need) {
    bases.resize(std::max(need, 2 * bases.size()));
    s.p = bases.data();
    s.cap = bases.size();
  };
  size_t shown = 0;
  for (int call = 0; rf.next_flat(chunk, threads, recs, text, sink); call++) {
    for (; shown < recs.size(); shown++) {
      const k2host::FlatRec &f = recs[shown];
      const char *t = text.data() + f.text_at;
      printf("%d|%.*s|%.*s|%.*s|%.*s\n", f.fastq ? 1 : 0, (int) f.id_len, t, (int) f.cmt_len, t + f.id_len,
             (int) f.seq_len, sink.p + f.seq_at, f.fastq ? (int) f.seq_len : 0, t + f.id_len + f.cmt_len);
    }
    if (call % 2) {
      recs.clear();
      text.clear();
      sink.n = 0;
      shown = 0;
    }
  }
  printf("----\n");
  // next_inplace: records indexed where they lie, the chunks copied as they
  // stand (cleared every other call)
  k2host::Reader ri;
  if (!ri.open(argv[1])) return 66;
  ri.set_pool(&workers);
  recs.clear();
  text.clear();
  std::vector<char> raw;
  k2host::SeqSink rs;
  rs.grow = [&](k2host::SeqSink &s, size_t need) {
    raw.resize(std::max(need, 2 * raw.size()));
    s.p = raw.data();
    s.cap = raw.size();
  };
  shown = 0;
  for (int call = 0; ri.next_inplace(chunk, threads, recs, text, rs); call++) {
    for (; shown < recs.size(); shown++) {
      const k2host::FlatRec &f = recs[shown];
      const char *t = text.data() + f.text_at;
      printf("%d|%.*s|%.*s|%.*s|%.*s\n", f.fastq ? 1 : 0, (int) f.id_len, t, (int) f.cmt_len, t + f.id_len,
             (int) f.seq_len, rs.p + f.seq_at, f.fastq ? (int) f.seq_len : 0, t + f.id_len + f.cmt_len);
    }
    if (call % 2) {
      recs.clear();
      text.clear();
      rs.n = 0;
      shown = 0;
    }
  }
  return 0;
}
