// FASTA/FASTQ input with the record rules of the reference's reader
// (src/seqreader.h:65-194 over klib kseq): a record starts at '>' or '@',
// the id runs to the first whitespace, the rest of the header line is the
// comment, sequence lines are joined until a line starting with '>', '+'
// or '@', and a '+' switches to reading as many quality characters as there
// are bases.  A line's single trailing '\r' is dropped.
#include <algorithm>
#include <cctype>
#include <chrono>
#include <cstring>
#include <thread>
#include <cerrno>
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include "k2g_host.h"

namespace k2host {

bool Reader::open(const char *path) {
  if (path) {
    fd_ = ::open(path, O_RDONLY);
    own_ = true;
  } else {
    fd_ = 0;
    own_ = false;
  }
  buf_.resize(1 << 20, 0);
  return fd_ >= 0;
}

Reader::~Reader() {
  if (own_ && fd_ > 2) ::close(fd_);
}

int Reader::get() {
  if (pos_ >= end_) {
    if (eof_) return -1;
    ssize_t n;
    do {
      n = ::read(fd_, buf_.data(), buf_.size());
    } while (n < 0 && errno == EINTR);
    if (n <= 0) {
      eof_ = true;
      return -1;
    }
    pos_ = 0;
    end_ = (size_t) n;
  }
  return (unsigned char) base()[pos_++];
}

// Rest of the current line into dst.
void Reader::line_into(std::string &dst, bool append, int *last) {
  if (!append) dst.clear();
  int c;
  for (;;) {
    // fast path: copy up to the next newline inside the buffer
    if (pos_ < end_) {
      const char *b = base() + pos_;
      const char *e = base() + end_;
      const char *nl = (const char *) memchr(b, '\n', e - b);
      if (nl) {
        dst.append(b, nl - b);
        pos_ = nl - base() + 1;
        c = '\n';
        break;
      }
      dst.append(b, e - b);
      pos_ = end_;
    }
    c = get();
    if (c < 0 || c == '\n') break;
    dst.push_back((char) c);
  }
  if (dst.size() > 1 && dst.back() == '\r') dst.pop_back();
  if (last) *last = c;
}

bool Reader::next(Record &r) {
  int c;
  if (pending_ == 0) {
    while ((c = get()) >= 0 && c != '>' && c != '@') {
    }
    if (c < 0) return false;
    pending_ = c;
  }
  r.header.clear();
  r.comment.clear();
  r.seq.clear();
  r.quals.clear();
  bool any = false;
  while ((c = get()) >= 0 && !isspace(c)) {
    r.header.push_back((char) c);
    any = true;
  }
  if (!any && c < 0) return false;
  if (c >= 0 && c != '\n') line_into(r.comment, false, nullptr);
  while ((c = get()) >= 0 && c != '>' && c != '+' && c != '@') {
    if (c == '\n') continue;
    r.seq.push_back((char) c);
    line_into(r.seq, true, nullptr);
  }
  if (c == '>' || c == '@') pending_ = c;
  if (c != '+') {
    r.fastq = false;
    if (c < 0) pending_ = 0;
    return true;
  }
  while ((c = get()) >= 0 && c != '\n') {
  }
  if (c < 0) return false;
  int last = 0;
  do {
    line_into(r.quals, true, &last);
  } while (last >= 0 && r.quals.size() < r.seq.size());
  pending_ = 0;
  if (r.quals.size() != r.seq.size()) return false;
  r.fastq = true;
  return true;
}

size_t Reader::load_block(size_t bases, std::vector<Record> &out) {
  size_t total = 0, n = 0;
  while (total < bases) {
    if (n == out.size()) out.emplace_back();
    if (!next(out[n])) break;
    total += out[n].seq.size();
    n++;
  }
  out.resize(n);
  return n;
}

size_t Reader::load_batch(size_t records, std::vector<Record> &out) {
  size_t n = 0;
  if (out.size() < records) out.resize(records);
  for (; n < records; n++)
    if (!next(out[n])) break;
  out.resize(n);
  return n;
}

void Record::append_text(std::string &out, const char *suffix) const {
  out.push_back(fastq ? '@' : '>');
  out += header;
  if (suffix) out += suffix;
  if (!comment.empty()) {
    out.push_back(' ');
    out += comment;
  }
  out.push_back('\n');
  out += seq;
  if (fastq) {
    out += "\n+\n";
    out += quals;
  }
  out.push_back('\n');
}

}  // namespace k2host

namespace k2host {

WorkPool::WorkPool(int threads) {
  for (int i = 1; i < threads; i++) workers_.emplace_back([this] { work(); });
}

WorkPool::~WorkPool() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  for (auto &t : workers_) t.join();
}

void WorkPool::work() {
  uint64_t seen = 0;
  for (;;) {
    const std::function<void(size_t)> *fn;
    size_t n;
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
      if (stop_) return;
      seen = gen_;
      fn = fn_;
      n = n_;
    }
    for (size_t i; (i = next_.fetch_add(1, std::memory_order_relaxed)) < n;) (*fn)(i);
    std::lock_guard<std::mutex> lk(mu_);
    if (++acked_ == workers_.size()) done_cv_.notify_all();
  }
}

void WorkPool::run(size_t n, const std::function<void(size_t)> &fn) {
  if (n == 0) return;
  if (workers_.empty() || n == 1) {
    for (size_t i = 0; i < n; i++) fn(i);
    return;
  }
  {
    std::lock_guard<std::mutex> lk(mu_);
    fn_ = &fn;
    n_ = n;
    next_.store(0, std::memory_order_relaxed);
    acked_ = 0;
    gen_++;
  }
  cv_.notify_all();
  for (size_t i; (i = next_.fetch_add(1, std::memory_order_relaxed)) < n;) fn(i);
  // every worker takes part in every run (if only to find no items left), so
  // none can still hold this run's fn when the next run starts
  std::unique_lock<std::mutex> lk(mu_);
  done_cv_.wait(lk, [&] { return acked_ == workers_.size(); });
}

void Reader::run(int t, const std::function<void(size_t)> &fn) {
  if (pool_) {
    pool_->run((size_t) t, fn);
    return;
  }
  std::vector<std::thread> ts;
  for (int i = 1; i < t; i++) ts.emplace_back(fn, (size_t) i);
  fn(0);
  for (auto &th : ts) th.join();
}

void Reader::open_memory(const char *p, size_t n) {
  fd_ = -1;
  own_ = false;
  eof_ = true;
  buf_.clear();
  ext_ = p;  // read in place
  pos_ = 0;
  end_ = n;
  pending_ = 0;
}

static double now_seconds() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// First record start ('>' right after '\n') at or after i in d[0..n), or n.
static size_t record_start(const char *d, size_t n, size_t i) {
  while (i < n) {
    const char *q = (const char *) memchr(d + i, '>', n - i);
    if (!q) return n;
    const size_t k = q - d;
    if (k > 0 && d[k - 1] == '\n') return k;
    i = k + 1;
  }
  return n;
}

// True if a line in d[0..n) starts with '@' or '+' (lines that end or
// switch a record in the record rules above).
static bool has_marker_line(const char *d, size_t n) {
  for (size_t i = 1; i < n; i++) {
    const char *q = (const char *) memchr(d + i, '\n', n - i);
    if (!q) return false;
    i = q - d + 1;
    if (i < n && (d[i] == '@' || d[i] == '+')) return true;
    i--;
  }
  return false;
}

size_t Reader::read_into(char *dst, size_t bytes, int threads) {
  size_t have = 0;
  struct stat st;
  if (!eof_ && fstat(fd_, &st) == 0 && S_ISREG(st.st_mode)) {
    // a regular file: copied in by several preads side by side, then the
    // descriptor's offset is moved past what was read
    const off_t at = lseek(fd_, 0, SEEK_CUR);
    if (at >= 0) {
      const size_t want = std::min<size_t>(bytes, st.st_size > at ? (size_t) (st.st_size - at) : 0);
      const int t = (int) std::max<size_t>(1, std::min<size_t>((size_t) threads, want >> 20));
      std::vector<ssize_t> got(t, 0);
      auto piece = [&](size_t i) {
        const size_t lo = want * i / t, hi = want * (i + 1) / t;
        size_t done = 0;
        while (lo + done < hi) {
          const ssize_t n = pread(fd_, dst + lo + done, hi - lo - done, at + (off_t) (lo + done));
          if (n < 0 && errno == EINTR) continue;
          if (n <= 0) break;
          done += (size_t) n;
        }
        got[i] = (ssize_t) done;
      };
      run(t, piece);
      bool whole = true;
      for (int i = 0; i < t && whole; i++) {
        const size_t lo = want * i / t, hi = want * (i + 1) / t;
        have += (size_t) got[i];
        whole = (size_t) got[i] == hi - lo;
      }
      lseek(fd_, at + (off_t) have, SEEK_SET);
      if (!whole) eof_ = true;  // the file ended (or shrank) inside the chunk
    }
  }
  while (have < bytes && !eof_) {
    ssize_t n;
    do {
      n = ::read(fd_, dst + have, bytes - have);
    } while (n < 0 && errno == EINTR);
    if (n <= 0)
      eof_ = true;
    else
      have += (size_t) n;
  }
  return have;
}

// The lines of a 4-line FASTQ record: header, one sequence line, '+' line,
// one quality line (line starts, ends before '\n', and the record's end).
struct Fastq4 {
  size_t h1, s0, s1, q0, end;
  uint64_t slen;
};

// True if d[p] starts a 4-line FASTQ record whose lines end inside d[0, n)
// (the quality line may end at n): '@', then a sequence line that next()
// reads as the whole sequence (not empty, not starting with '>', '+' or
// '@'), a line starting with '+', and a quality line of as many characters
// as bases (a line's last '\r' dropped when the text is longer than one
// character).  next() reads such a record the same way; anything else is
// left to it.
static bool fastq4(const char *d, size_t n, size_t p, Fastq4 &r) {
  if (p >= n || d[p] != '@') return false;
  auto line_end = [&](size_t i) -> size_t {
    const char *nl = i < n ? (const char *) memchr(d + i, '\n', n - i) : nullptr;
    return nl ? (size_t) (nl - d) : n;
  };
  r.h1 = line_end(p);
  r.s0 = r.h1 + 1;
  if (r.s0 >= n) return false;
  const char c = d[r.s0];
  if (c == '\n' || c == '>' || c == '+' || c == '@') return false;
  r.s1 = line_end(r.s0);
  const size_t p0 = r.s1 + 1;
  if (p0 >= n || d[p0] != '+') return false;
  const size_t p1 = line_end(p0);
  if (p1 >= n) return false;
  r.q0 = p1 + 1;
  const size_t q1 = line_end(r.q0);
  uint64_t slen = r.s1 - r.s0, qlen = q1 - r.q0;
  if (slen > 1 && d[r.s1 - 1] == '\r') slen--;
  if (qlen > 1 && d[q1 - 1] == '\r') qlen--;
  if (qlen != slen) return false;
  r.slen = slen;
  r.end = q1 < n ? q1 + 1 : n;
  return true;
}

// First 4-line FASTQ record start at or after i in d[0, n), or n.
static size_t fastq_record_start(const char *d, size_t n, size_t i) {
  Fastq4 r;
  if (i == 0) return fastq4(d, n, 0, r) ? 0 : n;
  while (i < n) {
    const char *nl = (const char *) memchr(d + i - 1, '\n', n - (i - 1));
    if (!nl) return n;
    const size_t k = (size_t) (nl - d) + 1;
    if (k < n && d[k] == '@' && fastq4(d, n, k, r)) return k;
    i = k + 1;
  }
  return n;
}

int Reader::fill_chunk(size_t bytes, int threads, size_t &cut) {
  // gather at least `bytes` unread bytes at the front of buf_
  if (pos_) {
    memmove(buf_.data(), buf_.data() + pos_, end_ - pos_);
    end_ -= pos_;
    pos_ = 0;
  }
  if (buf_.size() < bytes + (1 << 20)) buf_.resize(bytes + (1 << 20), end_);
  if (end_ < bytes && !eof_) end_ += read_into(buf_.data() + end_, bytes - end_, threads);
  const char *d = buf_.data();
  const size_t n = end_;
  if (n == 0) return 0;
  if (d[0] == '@') {
    // FASTQ: cut after the last whole 4-line record (everything, at the end
    // of input; the parts' parse checks that the records tile the chunk)
    cut = eof_ ? n : 0;
    Fastq4 r;
    for (size_t k = n; !eof_ && k-- > 1;)
      if (d[k] == '@' && d[k - 1] == '\n' && fastq4(d, n, k, r) && r.end < n) {
        cut = r.end;
        break;
      }
    if (!eof_ && cut == 0 && fastq4(d, n, 0, r) && r.end < n) cut = r.end;
    if (cut == 0) {
      serial_ = true;
      return 2;
    }
    return 3;
  }
  // cut after the last complete record (everything, at the end of input)
  cut = n;
  if (!eof_) {
    cut = 0;
    for (size_t k = n; k-- > 1;)
      if (d[k] == '>' && d[k - 1] == '\n') {
        cut = k;
        break;
      }
  }
  if (d[0] != '>' || cut == 0) {
    serial_ = true;  // not plain FASTA here: the record rules need one reader
    return 2;
  }
  return 1;
}

// Part boundaries of [0, cut): t parts starting at record starts.
static std::vector<size_t> split_parts(const char *d, size_t cut, int t, bool fastq = false) {
  std::vector<size_t> b(t + 1);
  b[0] = 0;
  for (int i = 1; i < t; i++)
    b[i] = std::max(b[i - 1], fastq ? fastq_record_start(d, cut, cut * i / t) : record_start(d, cut, cut * i / t));
  b[t] = cut;
  return b;
}

// Records of a FASTQ part d[0, n) (starting at a record start), indexed
// where they lie: every record must be a 4-line record (fastq4) and the
// records must end exactly at n, else `unusual` is set and the chunk is left
// to next().  Text per record: id, comment, qualities.
static void index_fastq(const char *d, size_t n, uint64_t at, std::vector<FlatRec> &recs, std::vector<char> &text,
                        bool quals, bool &unusual) {
  size_t i = 0;
  Fastq4 q;
  while (i < n) {
    if (!fastq4(d, n, i, q) || q.end > n) {
      unusual = true;
      return;
    }
    size_t j = i + 1;
    while (j < q.h1 && !isspace((unsigned char) d[j])) j++;
    const size_t id_len = j - (i + 1);
    size_t c0 = j, cmt_len = 0;
    if (j < q.h1) {  // the whitespace after the id, then the comment
      c0 = j + 1;
      cmt_len = q.h1 - c0;
      if (cmt_len > 1 && d[q.h1 - 1] == '\r') cmt_len--;
    }
    FlatRec r;
    r.text_at = text.size();
    r.id_len = (uint32_t) id_len;
    r.cmt_len = (uint32_t) cmt_len;
    r.fastq = true;
    r.seq_at = at + q.s0;
    r.seq_len = q.slen;
    text.insert(text.end(), d + i + 1, d + i + 1 + id_len);
    text.insert(text.end(), d + c0, d + c0 + cmt_len);
    if (quals) text.insert(text.end(), d + q.q0, d + q.q0 + q.slen);
    recs.push_back(r);
    i = q.end;
  }
}

size_t Reader::next_records(size_t bytes, int threads, std::vector<Record> &out, std::vector<Record> *pool) {
  const size_t before = out.size();
  auto serial = [&] {
    size_t got = 0;
    Record r;
    while (got < bytes && next(r)) {
      got += r.header.size() + r.seq.size() + r.quals.size() + 3;
      out.push_back(std::move(r));
      r = Record();
    }
    return out.size() - before;
  };
  if (serial_ || threads <= 1 || pending_) {
    serial_ = true;
    return serial();
  }
  size_t cut = 0;
  const int got = fill_chunk(bytes, threads, cut);
  if (got == 0) return 0;
  if (got != 1) {  // FASTQ is read by next() here
    serial_ = true;
    return serial();
  }
  const char *d = buf_.data();
  // split [0, cut) at record starts and parse the parts side by side
  const int t = (int) std::min<size_t>((size_t) threads, std::max<size_t>(1, cut >> 12));
  const std::vector<size_t> b = split_parts(d, cut, t);
  std::vector<std::vector<Record>> parts(t), spare(t);
  if (pool) {  // hand each part a share of the used records
    const size_t share = pool->size() / t;
    for (int i = 0; i < t; i++)
      for (size_t k = 0; k < share; k++) {
        spare[i].push_back(std::move(pool->back()));
        pool->pop_back();
      }
  }
  std::vector<char> marker(t, 0);
  auto parse = [&](size_t i) {
    if (b[i + 1] <= b[i]) return;
    // every line start of the chunk lies in exactly one part (parts begin at
    // record starts)
    if (has_marker_line(d + b[i], b[i + 1] - b[i])) {
      marker[i] = 1;
      return;
    }
    Reader r;
    r.open_memory(d + b[i], b[i + 1] - b[i]);
    for (;;) {
      Record rec;
      if (!spare[i].empty()) {
        rec = std::move(spare[i].back());
        spare[i].pop_back();
      }
      if (!r.next(rec)) {
        spare[i].push_back(std::move(rec));
        break;
      }
      parts[i].push_back(std::move(rec));
    }
  };
  run(t, parse);
  if (std::find(marker.begin(), marker.end(), 1) != marker.end()) {
    // a line starts with '@' or '+': the record rules need one reader from
    // this chunk on (the parsed parts are dropped, their buffers kept)
    if (pool)
      for (int i = 0; i < t; i++) {
        for (auto &rec : parts[i]) pool->push_back(std::move(rec));
        for (auto &rec : spare[i]) pool->push_back(std::move(rec));
      }
    serial_ = true;
    return serial();
  }
  for (auto &pt : parts)
    for (auto &rec : pt) out.push_back(std::move(rec));
  if (pool)
    for (auto &sp : spare)
      for (auto &rec : sp) pool->push_back(std::move(rec));
  pos_ = cut;
  return out.size() - before;
}

// Records, text and bases of a plain-FASTA part d[0, n): n starts at a
// record start, and with WRITE false only counts them (and stops at a line
// starting with '@' or '+', setting marker).  The record rules of next()
// over whole lines: the id runs to the first whitespace, the rest of that
// line is the comment, sequence lines are joined up to the next line
// starting with '>', empty lines are skipped, and a line's last '\r' is
// dropped when the text it ends is longer than one character.
struct FlatCount {
  size_t recs = 0, text = 0, seq = 0;
  bool marker = false;
};

template <bool WRITE>
static void scan_fasta(const char *d, size_t n, FlatCount &c, FlatRec *recs, char *text, uint64_t text_at,
                       char *seq, uint64_t seq_at) {
  size_t i = 0;
  while (i < n) {
    i++;  // the '>' of the record
    const size_t id0 = i;
    while (i < n && !isspace((unsigned char) d[i])) i++;
    const size_t id_len = i - id0;
    if (id_len == 0 && i == n) break;  // a lone '>' at the end of input: no record
    size_t c0 = i, cmt_len = 0;
    if (i < n && d[i] != '\n') {
      c0 = i + 1;
      const char *nl = (const char *) memchr(d + c0, '\n', n - c0);
      const size_t le = nl ? (size_t) (nl - d) : n;
      cmt_len = le - c0;
      if (cmt_len > 1 && d[le - 1] == '\r') cmt_len--;
      i = nl ? le + 1 : n;
    } else if (i < n) {
      i++;
    }
    char *sp = WRITE ? seq + c.seq : nullptr;
    size_t slen = 0;
    while (i < n && d[i] != '>') {
      if (d[i] == '\n') {
        i++;
        continue;
      }
      if (!WRITE && (d[i] == '@' || d[i] == '+')) {
        c.marker = true;
        return;
      }
      const char *nl = (const char *) memchr(d + i, '\n', n - i);
      const size_t le = nl ? (size_t) (nl - d) : n;
      // the dropped '\r' is not written: the byte after a part's last base
      // belongs to the next part
      size_t keep = le - i;
      if (slen + keep > 1 && d[le - 1] == '\r') keep--;
      if (WRITE) memcpy(sp + slen, d + i, keep);
      slen += keep;
      i = nl ? le + 1 : n;
    }
    if (WRITE) {
      FlatRec &r = recs[c.recs];
      r.text_at = text_at + c.text;
      r.id_len = (uint32_t) id_len;
      r.cmt_len = (uint32_t) cmt_len;
      r.seq_at = seq_at + c.seq;
      r.seq_len = slen;
      r.fastq = false;
      memcpy(text + c.text, d + id0, id_len);
      memcpy(text + c.text + id_len, d + c0, cmt_len);
    }
    c.recs++;
    c.text += id_len + cmt_len;
    c.seq += slen;
  }
}

static void append_flat(const Record &r, std::vector<FlatRec> &recs, std::vector<char> &text, SeqSink &seq,
                        bool quals = true) {
  FlatRec f;
  f.text_at = text.size();
  f.id_len = (uint32_t) r.header.size();
  f.cmt_len = (uint32_t) r.comment.size();
  f.fastq = r.fastq;
  text.insert(text.end(), r.header.begin(), r.header.end());
  text.insert(text.end(), r.comment.begin(), r.comment.end());
  if (r.fastq && quals) text.insert(text.end(), r.quals.begin(), r.quals.end());  // as many as bases
  f.seq_at = seq.n;
  f.seq_len = r.seq.size();
  if (f.seq_len) memcpy(seq.room(f.seq_len), r.seq.data(), f.seq_len);
  seq.n += f.seq_len;
  recs.push_back(f);
}

// Records of a plain-FASTA part d[0, n) indexed where they lie (the record
// rules of scan_fasta): seq_at = at + the offset of the record's first
// sequence line, seq_len its bases.  Records whose bases span several lines
// are listed in `joins` (record, end of its text) for join_lines; nothing
// is written to d.  Stops at a line starting with '@' or '+' (marker).
static void index_fasta(const char *d, size_t n, uint64_t at, std::vector<FlatRec> &recs, std::vector<char> &text,
                        std::vector<InplaceJoin> &joins, bool &marker) {
  size_t i = 0;
  while (i < n) {
    i++;  // the '>' of the record
    const size_t id0 = i;
    while (i < n && !isspace((unsigned char) d[i])) i++;
    const size_t id_len = i - id0;
    if (id_len == 0 && i == n) break;  // a lone '>' at the end of input: no record
    size_t c0 = i, cmt_len = 0;
    if (i < n && d[i] != '\n') {
      c0 = i + 1;
      const char *nl = (const char *) memchr(d + c0, '\n', n - c0);
      const size_t le = nl ? (size_t) (nl - d) : n;
      cmt_len = le - c0;
      if (cmt_len > 1 && d[le - 1] == '\r') cmt_len--;
      i = nl ? le + 1 : n;
    } else if (i < n) {
      i++;
    }
    FlatRec r;
    r.text_at = text.size();
    r.id_len = (uint32_t) id_len;
    r.cmt_len = (uint32_t) cmt_len;
    text.insert(text.end(), d + id0, d + id0 + id_len);
    text.insert(text.end(), d + c0, d + c0 + cmt_len);
    uint64_t slen = 0;
    size_t lines = 0, first = i;
    while (i < n && d[i] != '>') {
      if (d[i] == '\n') {
        i++;
        continue;
      }
      if (d[i] == '@' || d[i] == '+') {
        marker = true;
        return;
      }
      const char *nl = (const char *) memchr(d + i, '\n', n - i);
      const size_t le = nl ? (size_t) (nl - d) : n;
      size_t keep = le - i;
      if (slen + keep > 1 && d[le - 1] == '\r') keep--;
      if (lines++ == 0) first = i;
      slen += keep;
      i = nl ? le + 1 : n;
    }
    r.seq_at = at + first;
    r.seq_len = slen;
    if (lines > 1) joins.push_back({recs.size(), at + i});
    recs.push_back(r);
  }
}

// Moves the sequence lines of a record that spans several lines up against
// its first line: raw[from, end) is the record's text from its first
// sequence line on (the rules of index_fasta again).
static void join_lines(char *raw, uint64_t from, uint64_t end) {
  uint64_t i = from, w = from, slen = 0;
  while (i < end) {
    if (raw[i] == '\n') {
      i++;
      continue;
    }
    const char *nl = (const char *) memchr(raw + i, '\n', end - i);
    const uint64_t le = nl ? (uint64_t) (nl - raw) : end;
    uint64_t keep = le - i;
    if (slen + keep > 1 && raw[le - 1] == '\r') keep--;
    if (w != i) memmove(raw + w, raw + i, keep);
    w += keep;
    slen += keep;
    i = nl ? le + 1 : end;
  }
}

size_t Reader::next_inplace(size_t bytes, int threads, std::vector<FlatRec> &recs, std::vector<char> &text,
                            SeqSink &raw) {
  const size_t before = recs.size();
  auto serial = [&] {
    size_t got = 0;
    Record r;
    while (got < bytes && next(r)) {
      got += r.header.size() + r.seq.size() + r.quals.size() + 3;
      append_flat(r, recs, text, raw, keep_quals_);
    }
    return recs.size() - before;
  };
  if (serial_ || threads <= 1 || pending_) {
    serial_ = true;
    return serial();
  }
  size_t cut = 0;
  const double t0 = now_seconds();
  const int got = fill_chunk(bytes, threads, cut);
  timing.fill += now_seconds() - t0;
  if (got == 0) return 0;
  if (got == 2) return serial();
  // FASTQ: the bases copied back to back (its quality lines are as long as
  // the bases, so the file's layout would double the bytes moved)
  if (got == 3) return flat_fastq(cut, threads, recs, text, raw, before) ? recs.size() - before : serial();
  char *d = buf_.data();
  const uint64_t at = raw.n;  // where d[0] lands in raw
  // parts of about 512 KiB: each is indexed and then copied while it is
  // still in the core's cache
  const int t = (int) std::max<size_t>(1, std::min<size_t>(cut >> 12, std::max<size_t>(threads, cut >> 19)));
  const std::vector<size_t> b = split_parts(d, cut, t);
  char *dst = raw.room(cut);
  // per-part arrays, kept across calls (their memory is reused)
  auto &prec = inplace_recs_;
  auto &ptext = inplace_text_;
  auto &pjoin = inplace_joins_;
  if ((int) prec.size() < t) {
    prec.resize(t);
    ptext.resize(t);
    pjoin.resize(t);
  }
  for (int i = 0; i < t; i++) {
    prec[i].clear();
    ptext[i].clear();
    pjoin[i].clear();
  }
  std::vector<char> marker(t, 0);
  double t1 = now_seconds();
  run(t, [&](size_t i) {
    bool m = false;
    index_fasta(d + b[i], b[i + 1] - b[i], at + b[i], prec[i], ptext[i], pjoin[i], m);
    marker[i] = m;
    if (!m) memcpy(dst + b[i], d + b[i], b[i + 1] - b[i]);
  });
  timing.count += now_seconds() - t1;
  if (std::find(marker.begin(), marker.end(), 1) != marker.end()) {
    serial_ = true;  // raw.n is unchanged: next() reads the chunk again
    return serial();
  }
  // multi-line records joined in raw, the parts' arrays one after another
  t1 = now_seconds();
  std::vector<size_t> rat(t + 1), tat(t + 1);
  rat[0] = before;
  tat[0] = text.size();
  for (int i = 0; i < t; i++) {
    rat[i + 1] = rat[i] + prec[i].size();
    tat[i + 1] = tat[i] + ptext[i].size();
  }
  recs.resize(rat[t]);
  text.resize(tat[t]);
  run(t, [&](size_t i) {
    for (const InplaceJoin &j : pjoin[i]) join_lines(raw.p, prec[i][j.rec].seq_at, j.end);
    for (size_t k = 0; k < prec[i].size(); k++) {
      FlatRec r = prec[i][k];
      r.text_at += tat[i];
      recs[rat[i] + k] = r;
    }
    if (!ptext[i].empty()) memcpy(text.data() + tat[i], ptext[i].data(), ptext[i].size());
  });
  raw.n += cut;
  timing.write += now_seconds() - t1;
  pos_ = cut;
  return recs.size() - before;
}

bool Reader::flat_fastq(size_t cut, int threads, std::vector<FlatRec> &recs, std::vector<char> &text, SeqSink &seq,
                        size_t before) {
  // the parts indexed, then each record's bases copied back to back
  const char *d = buf_.data();
  const int t = (int) std::max<size_t>(1, std::min<size_t>(cut >> 12, std::max<size_t>(threads, cut >> 19)));
  const std::vector<size_t> b = split_parts(d, cut, t, true);
  auto &prec = inplace_recs_;
  auto &ptext = inplace_text_;
  if ((int) prec.size() < t) {
    prec.resize(t);
    ptext.resize(t);
    inplace_joins_.resize(t);
  }
  std::vector<char> unusual(t, 0);
  double t1 = now_seconds();
  run(t, [&](size_t i) {
    prec[i].clear();
    ptext[i].clear();
    bool m = false;
    index_fastq(d + b[i], b[i + 1] - b[i], b[i], prec[i], ptext[i], keep_quals_, m);
    unusual[i] = m;
  });
  timing.count += now_seconds() - t1;
  if (std::find(unusual.begin(), unusual.end(), 1) != unusual.end()) {
    serial_ = true;  // nothing was written: next() reads the chunk again
    return false;
  }
  t1 = now_seconds();
  std::vector<uint64_t> rat(t + 1), tat(t + 1), sat(t + 1);
  rat[0] = before;
  tat[0] = text.size();
  sat[0] = seq.n;
  for (int i = 0; i < t; i++) {
    uint64_t bases = 0;
    for (const FlatRec &r : prec[i]) bases += r.seq_len;
    rat[i + 1] = rat[i] + prec[i].size();
    tat[i + 1] = tat[i] + ptext[i].size();
    sat[i + 1] = sat[i] + bases;
  }
  recs.resize(rat[t]);
  text.resize(tat[t]);
  seq.room(sat[t] - seq.n);
  run(t, [&](size_t i) {
    uint64_t pos = sat[i];
    for (size_t k = 0; k < prec[i].size(); k++) {
      FlatRec r = prec[i][k];
      memcpy(seq.p + pos, d + r.seq_at, r.seq_len);
      r.seq_at = pos;
      r.text_at += tat[i];
      pos += r.seq_len;
      recs[rat[i] + k] = r;
    }
    if (!ptext[i].empty()) memcpy(text.data() + tat[i], ptext[i].data(), ptext[i].size());
  });
  seq.n = sat[t];
  timing.write += now_seconds() - t1;
  pos_ = cut;
  return true;
}

size_t Reader::next_flat(size_t bytes, int threads, std::vector<FlatRec> &recs, std::vector<char> &text,
                         SeqSink &seq) {
  const size_t before = recs.size();
  auto serial = [&] {
    size_t got = 0;
    Record r;
    while (got < bytes && next(r)) {
      got += r.header.size() + r.seq.size() + r.quals.size() + 3;
      append_flat(r, recs, text, seq, keep_quals_);
    }
    return recs.size() - before;
  };
  if (serial_ || threads <= 1 || pending_) {
    serial_ = true;
    return serial();
  }
  size_t cut = 0;
  const double t0 = now_seconds();
  const int got = fill_chunk(bytes, threads, cut);
  timing.fill += now_seconds() - t0;
  if (got == 0) return 0;
  if (got == 2) return serial();
  const char *d = buf_.data();
  if (got == 3) return flat_fastq(cut, threads, recs, text, seq, before) ? recs.size() - before : serial();
  const int t = (int) std::min<size_t>((size_t) threads, std::max<size_t>(1, cut >> 12));
  const std::vector<size_t> b = split_parts(d, cut, t);
  std::vector<FlatCount> cnt(t);
  double t1 = now_seconds();
  run(t, [&](size_t i) { scan_fasta<false>(d + b[i], b[i + 1] - b[i], cnt[i], nullptr, nullptr, 0, nullptr, 0); });
  double t2 = now_seconds();
  timing.count += t2 - t1;
  for (int i = 0; i < t; i++)
    if (cnt[i].marker) {
      // a line starts with '@' or '+': the record rules need one reader from
      // this chunk on
      serial_ = true;
      return serial();
    }
  std::vector<FlatCount> at(t);  // where each part writes
  FlatCount total;
  total.recs = before;
  total.text = text.size();
  total.seq = seq.n;
  for (int i = 0; i < t; i++) {
    at[i] = total;
    total.recs += cnt[i].recs;
    total.text += cnt[i].text;
    total.seq += cnt[i].seq;
  }
  recs.resize(total.recs);
  text.resize(total.text);
  seq.room(total.seq - seq.n);
  t1 = now_seconds();
  run(t, [&](size_t i) {
    FlatCount c;
    scan_fasta<true>(d + b[i], b[i + 1] - b[i], c, recs.data() + at[i].recs, text.data() + at[i].text,
                     at[i].text, seq.p + at[i].seq, at[i].seq);
  });
  timing.write += now_seconds() - t1;
  seq.n = total.seq;
  pos_ = cut;
  return recs.size() - before;
}

}  // namespace k2host
