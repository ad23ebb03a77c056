// FASTA/FASTQ input with the record rules of the reference's reader
// (src/seqreader.h:65-194 over klib kseq): a record starts at '>' or '@',
// the id runs to the first whitespace, the rest of the header line is the
// comment, sequence lines are joined until a line starting with '>', '+'
// or '@', and a '+' switches to reading as many quality characters as there
// are bases.  A line's single trailing '\r' is dropped.
#include <cctype>
#include <cstring>
#include <cerrno>
#include <fcntl.h>
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
  buf_.resize(1 << 20);
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
  return (unsigned char) buf_[pos_++];
}

// Rest of the current line into dst.
void Reader::line_into(std::string &dst, bool append, int *last) {
  if (!append) dst.clear();
  int c;
  for (;;) {
    // fast path: copy up to the next newline inside the buffer
    if (pos_ < end_) {
      const char *b = buf_.data() + pos_;
      const char *e = buf_.data() + end_;
      const char *nl = (const char *) memchr(b, '\n', e - b);
      if (nl) {
        dst.append(b, nl - b);
        pos_ = nl - buf_.data() + 1;
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
