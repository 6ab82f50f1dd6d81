/*
 * Small I/O helpers of the eqtlbma_bf host program: gzip line reading and
 * writing through zlib (plain text is read transparently), tokenising with
 * the reference's rules (strtok semantics: runs of delimiters collapse,
 * utils_io.cpp:46-62) and the number formats of its writers.
 */
#ifndef EQBF_HOST_UTIL_HPP
#define EQBF_HOST_UTIL_HPP

#include <zlib.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace host {

/* Prints "ERROR: msg" on stderr and exits with EXIT_FAILURE, as the
 * reference does for every fatal condition (utils_io.cpp:214-219). */
[[noreturn]] void fail(const std::string& msg);

/* Reads a (possibly gzipped) text file line by line; '\n' ends a line and
 * is not kept, and a last line without '\n' is dropped, as the reference's
 * reader does (utils_io.cpp:268-285). */
class LineReader {
 public:
  explicit LineReader(const std::string& path);
  ~LineReader();
  LineReader(const LineReader&) = delete;
  LineReader& operator=(const LineReader&) = delete;
  bool next(std::string& line);
  bool complete() const { return eof_; } /* the whole file was read */
  const std::string& path() const { return path_; }

 private:
  gzFile f_ = nullptr;
  std::string path_;
  bool eof_ = false;
  std::vector<char> buf_;
};

/* Appends to a gzip file: mode "wb" starts the file, "ab" adds a member. */
class GzWriter {
 public:
  GzWriter(const std::string& path, const char* mode);
  ~GzWriter();
  GzWriter(const GzWriter&) = delete;
  GzWriter& operator=(const GzWriter&) = delete;
  void put(const std::string& s);

 private:
  gzFile f_ = nullptr;
  std::string path_;
};

/* Splits s at any of the delimiter characters, dropping empty tokens. */
void tokenize(const std::string& s, const char* delims, std::vector<std::string>& out);

/* iostream scientific with precision 6 (the writers' format): "%.6e";
 * NaN prints "nan", infinities "inf" / "-inf". */
void put_sci(std::string& out, double v);
/* iostream default float format with precision 6 (like "%g"). */
void put_gen(std::string& out, double v);
void put_int(std::string& out, int64_t v);

bool file_exists(const std::string& path);

/* Whole-file read with the rules of the reference's readFile
 * (utils_io.cpp:290-376): the last byte of the file is dropped (normally
 * its final newline), lines split at '\n' with empty ones removed. */
std::vector<std::string> read_file_lines(const std::string& path);

}  // namespace host

#endif
