#include "util.hpp"

#include <sys/stat.h>

#include <cstdlib>
#include <cstring>
#include <iostream>

namespace host {

void fail(const std::string& msg) {
  std::cerr << "ERROR: " << msg << std::endl;
  std::exit(EXIT_FAILURE);
}

LineReader::LineReader(const std::string& path) : path_(path), buf_(1 << 16) {
  f_ = gzopen(path.c_str(), "rb");
  if (!f_) fail("can't open file " + path);
  gzbuffer(f_, 1 << 20);
}

LineReader::~LineReader() {
  if (f_) gzclose(f_);
}

bool LineReader::next(std::string& line) {
  line.clear();
  if (eof_) return false;
  while (true) {
    char* r = gzgets(f_, buf_.data(), (int)buf_.size());
    if (!r) {
      int err = 0;
      gzerror(f_, &err);
      if (err != Z_OK && err != Z_BUF_ERROR) fail("can't read successfully file " + path_ + " up to the end");
      eof_ = true;
      /* like the reference's reader, a last line without '\n' is not
       * returned (utils_io.cpp:273-285) */
      return false;
    }
    const size_t n = std::strlen(r);
    if (n > 0 && r[n - 1] == '\n') {
      line.append(r, n - 1);
      return true;
    }
    line.append(r, n);
  }
}

GzWriter::GzWriter(const std::string& path, const char* mode) : path_(path) {
  f_ = gzopen(path.c_str(), mode);
  if (!f_) fail("can't open file " + path);
}

GzWriter::~GzWriter() {
  if (f_ && gzclose(f_) != Z_OK) {
    std::cerr << "ERROR: can't close file " << path_ << std::endl;
    std::exit(EXIT_FAILURE);
  }
}

void GzWriter::put(const std::string& s) {
  if (s.empty()) return;
  if (gzwrite(f_, s.data(), (unsigned)s.size()) != (int)s.size()) fail("can't write to file " + path_);
}

void tokenize(const std::string& s, const char* delims, std::vector<std::string>& out) {
  out.clear();
  size_t i = 0;
  const size_t n = s.size();
  while (i < n) {
    while (i < n && std::strchr(delims, s[i]) && s[i] != '\0') ++i;
    if (i >= n) break;
    size_t j = i;
    while (j < n && !(std::strchr(delims, s[j]) && s[j] != '\0')) ++j;
    out.emplace_back(s, i, j - i);
    i = j;
  }
}

void put_sci(std::string& out, double v) {
  char b[64];
  if (std::isnan(v)) {
    out += "nan";
    return;
  }
  if (std::isinf(v)) {
    out += v > 0 ? "inf" : "-inf";
    return;
  }
  std::snprintf(b, sizeof b, "%.6e", v);
  out += b;
}

void put_gen(std::string& out, double v) {
  char b[64];
  if (std::isnan(v)) {
    out += "nan";
    return;
  }
  if (std::isinf(v)) {
    out += v > 0 ? "inf" : "-inf";
    return;
  }
  std::snprintf(b, sizeof b, "%g", v);
  out += b;
}

void put_int(std::string& out, int64_t v) { out += std::to_string(v); }

bool file_exists(const std::string& path) {
  struct stat st;
  return stat(path.c_str(), &st) == 0;
}

std::vector<std::string> read_file_lines(const std::string& path) {
  gzFile f = gzopen(path.c_str(), "rb");
  if (!f) fail("can't open file " + path);
  gzbuffer(f, 1 << 20);
  std::string buf;
  std::vector<char> chunk(1 << 20);
  while (true) {
    const int n = gzread(f, chunk.data(), (unsigned)chunk.size());
    if (n < 0) fail("gzread failed on " + path);
    if (n == 0) break;
    buf.append(chunk.data(), (size_t)n);
  }
  gzclose(f);
  if (!buf.empty()) buf.pop_back();
  std::vector<std::string> out;
  tokenize(buf, "\n", out);
  return out;
}

}  // namespace host
