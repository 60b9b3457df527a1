// Relation files: the text format the reference's generator writes and its
// join data builder reads (datagen/src/bin/generate.rs:343-365 writes a
// header record and one `key value` record per tuple, value = row + 1;
// data-store/src/join_data.rs:173-280 reads them with a space delimiter,
// the header skipped, gzip when the path ends in "gz", every record
// deserialized as two integers of the key type).
//
// Host code only: the columns are read into / written from host memory.
// gzopen reads plain files unchanged, so one path serves both.
#include <zlib.h>

#include <cerrno>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../../include/fij.h"

namespace fij {
fij_status set_error(fij_status s, const char* msg);
}

namespace {

fij_status io_fail(fij_status s, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return fij::set_error(s, buf);
}

bool ends_with_gz(const char* path) {
  const size_t n = strlen(path);
  return n >= 2 && path[n - 2] == 'g' && path[n - 1] == 'z';
}

// Streams the records of a relation file: each call to next() yields the
// fields of one non-empty line (CR before LF is dropped), the header first.
class RecordReader {
 public:
  // gzip is recognised by content, so a compressed file is read whatever
  // its name, and a plain one too
  explicit RecordReader(const char* path) {
    f_ = gzopen(path, "rb");
    if (f_) gzbuffer(f_, 1 << 20);
  }
  ~RecordReader() {
    if (f_) gzclose(f_);
  }
  bool ok() const { return f_ != nullptr; }

  // Returns 1 with the line in [b, e), 0 at the end, -1 on a read error.
  int next_line(const char*& b, const char*& e) {
    for (;;) {
      const char* nl = static_cast<const char*>(memchr(buf_.data() + pos_, '\n', len_ - pos_));
      if (nl) {
        b = buf_.data() + pos_;
        e = nl;
        pos_ = static_cast<size_t>(nl - buf_.data()) + 1;
        ++line_;
        if (e > b && e[-1] == '\r') --e;
        if (e == b) continue;  // blank lines are skipped
        return 1;
      }
      if (eof_) {
        if (pos_ == len_) return 0;
        b = buf_.data() + pos_;
        e = buf_.data() + len_;
        pos_ = len_;
        ++line_;
        if (e > b && e[-1] == '\r') --e;
        if (e == b) return 0;
        return 1;
      }
      // refill: keep the partial line
      if (pos_ > 0) {
        memmove(buf_.data(), buf_.data() + pos_, len_ - pos_);
        len_ -= pos_;
        pos_ = 0;
      }
      if (buf_.size() - len_ < (1u << 20)) buf_.resize(buf_.size() * 2 + (1u << 20));
      const int got = gzread(f_, buf_.data() + len_, static_cast<unsigned>(buf_.size() - len_));
      if (got < 0) return -1;
      if (got == 0) eof_ = true;
      len_ += static_cast<size_t>(got);
    }
  }
  uint64_t line() const { return line_; }
  const char* error() {
    int err = 0;
    const char* m = f_ ? gzerror(f_, &err) : nullptr;
    return m ? m : strerror(errno);
  }

 private:
  gzFile f_ = nullptr;
  std::vector<char> buf_ = std::vector<char>(1 << 20);
  size_t pos_ = 0, len_ = 0;
  bool eof_ = false;
  uint64_t line_ = 0;
};

// Parses one integer field as Rust's str::parse does: an optional sign and
// at least one decimal digit, nothing else, within the type's range.
template <typename T>
bool parse_int(const char* b, const char* e, T& out) {
  if (b == e) return false;
  bool neg = false;
  if (*b == '+' || *b == '-') {
    neg = *b == '-';
    if (++b == e) return false;
  }
  using U = unsigned long long;
  const U limit = neg ? static_cast<U>(std::numeric_limits<T>::max()) + 1 : static_cast<U>(std::numeric_limits<T>::max());
  U v = 0;
  for (; b < e; ++b) {
    const unsigned d = static_cast<unsigned>(*b - '0');
    if (d > 9) return false;
    if (v > (limit - d) / 10) return false;
    v = v * 10 + d;
  }
  out = neg ? static_cast<T>(0 - v) : static_cast<T>(v);
  return true;
}

template <typename T>
fij_status read_relation(const char* path, char delim, T* keys, T* pays, uint64_t capacity, uint64_t* len,
                         bool count_only) {
  RecordReader r(path);
  if (!r.ok()) return io_fail(FIJ_IO_ERROR, "cannot open %s: %s", path, strerror(errno));
  const char *b, *e;
  int st = r.next_line(b, e);  // header
  if (st < 0) return io_fail(FIJ_IO_ERROR, "%s: %s", path, r.error());
  if (st == 1) {  // records must have as many fields as the header: two
    const char* d = static_cast<const char*>(memchr(b, delim, static_cast<size_t>(e - b)));
    if (!d || memchr(d + 1, delim, static_cast<size_t>(e - d - 1)))
      return io_fail(FIJ_PARSE_ERROR, "%s:1: the header must name two fields separated by one '%c'", path, delim);
  }
  uint64_t n = 0;
  while (st == 1 && (st = r.next_line(b, e)) == 1) {
    if (count_only) {
      ++n;
      continue;
    }
    const char* d = static_cast<const char*>(memchr(b, delim, static_cast<size_t>(e - b)));
    T k{}, v{};
    if (!d || memchr(d + 1, delim, static_cast<size_t>(e - d - 1)) || !parse_int(b, d, k) || !parse_int(d + 1, e, v))
      return io_fail(FIJ_PARSE_ERROR, "%s:%llu: expected two integers separated by one '%c'", path,
                     static_cast<unsigned long long>(r.line()), delim);
    if (n >= capacity)
      return io_fail(FIJ_INVALID_ARGUMENT, "%s holds more than %llu tuples", path,
                     static_cast<unsigned long long>(capacity));
    keys[n] = k;
    pays[n] = v;
    ++n;
  }
  if (st < 0) return io_fail(FIJ_IO_ERROR, "%s: %s", path, r.error());
  *len = n;
  return FIJ_OK;
}

template <typename T>
fij_status write_relation(const char* path, char delim, const T* keys, const T* pays, uint64_t len) {
  const bool gz = ends_with_gz(path);
  gzFile f = gzopen(path, gz ? "wb6" : "wbT");  // T: plain, no compression
  if (!f) return io_fail(FIJ_IO_ERROR, "cannot create %s: %s", path, strerror(errno));
  gzbuffer(f, 1 << 20);
  std::string out;
  out.reserve(1 << 21);
  out += "key";
  out += delim;
  out += "value\n";
  char tmp[48];
  bool ok = true;
  for (uint64_t i = 0; i < len && ok; ++i) {
    const int m = snprintf(tmp, sizeof(tmp), "%lld%c%lld\n", static_cast<long long>(keys[i]), delim,
                           static_cast<long long>(pays[i]));
    out.append(tmp, static_cast<size_t>(m));
    if (out.size() >= (1u << 20)) {
      ok = gzwrite(f, out.data(), static_cast<unsigned>(out.size())) == static_cast<int>(out.size());
      out.clear();
    }
  }
  if (ok && !out.empty()) ok = gzwrite(f, out.data(), static_cast<unsigned>(out.size())) == static_cast<int>(out.size());
  const int c = gzclose(f);
  if (!ok || c != Z_OK) return io_fail(FIJ_IO_ERROR, "writing %s failed", path);
  return FIJ_OK;
}

char delimiter(char d) { return d ? d : ' '; }

}  // namespace

extern "C" {

fij_status fij_relation_file_len(const char* path, char delim, uint64_t* len) {
  if (!path || !len) return io_fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  return read_relation<int64_t>(path, delimiter(delim), nullptr, nullptr, 0, len, true);
}

fij_status fij_relation_file_read(const char* path, char delim, fij_type type, void* keys, void* payloads,
                                  uint64_t capacity, uint64_t* len) {
  if (!path || !len || (capacity && (!keys || !payloads)))
    return io_fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  if (type == FIJ_TYPE_I32)
    return read_relation<int32_t>(path, delimiter(delim), static_cast<int32_t*>(keys),
                                  static_cast<int32_t*>(payloads), capacity, len, false);
  if (type == FIJ_TYPE_I64)
    return read_relation<int64_t>(path, delimiter(delim), static_cast<int64_t*>(keys),
                                  static_cast<int64_t*>(payloads), capacity, len, false);
  return io_fail(FIJ_INVALID_ARGUMENT, "unknown type %d", static_cast<int>(type));
}

fij_status fij_relation_file_write(const char* path, char delim, fij_type type, const void* keys,
                                   const void* payloads, uint64_t len) {
  if (!path || (len && (!keys || !payloads))) return io_fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  if (type == FIJ_TYPE_I32)
    return write_relation<int32_t>(path, delimiter(delim), static_cast<const int32_t*>(keys),
                                   static_cast<const int32_t*>(payloads), len);
  if (type == FIJ_TYPE_I64)
    return write_relation<int64_t>(path, delimiter(delim), static_cast<const int64_t*>(keys),
                                   static_cast<const int64_t*>(payloads), len);
  return io_fail(FIJ_INVALID_ARGUMENT, "unknown type %d", static_cast<int>(type));
}

}  // extern "C"
