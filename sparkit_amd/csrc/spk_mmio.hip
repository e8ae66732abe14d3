// spk_mmio.hip -- Matrix Market text in and out of device stores
// (matrix_market.hpp:43-163, matrix_market.cpp:28-133).  The text is
// scanned on the host; the (row, col, value) lists it yields are assembled
// on the device by the CSR assembly (sort, last-of-run wins), which is
// to_compressed_row of the reference's Coordinate_matrix (insert_or_assign:
// the last entry at a position wins).
#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "spk_csr.hpp"

namespace spk {
namespace {

[[noreturn]] void runtime_fail(const std::string& msg) {
  throw Error{SPK_RUNTIME_ERROR, "matrix market: " + msg};
}

std::string lower(std::string s) {
  for (auto& c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

// std::getline over a memory buffer: false only when nothing is left
struct LineReader {
  const char* p;
  const char* end;
  bool next(std::string& line) {
    line.clear();
    if (p >= end) return false;
    const char* q = p;
    while (q < end && *q != '\n') ++q;
    line.assign(p, q);
    p = q < end ? q + 1 : q;
    return true;
  }
};

// parse_banner (matrix_market.cpp:28-95): returns {symmetric, pattern}
void parse_banner(const std::string& line, bool* symmetric, bool* pattern) {
  std::istringstream iss(line);
  std::string header, object, format, field, symmetry;
  iss >> header >> object >> format >> field >> symmetry;
  if (lower(header) != "%%matrixmarket") runtime_fail("invalid banner header: " + header);
  format = lower(format);
  field = lower(field);
  symmetry = lower(symmetry);
  if (format != "coordinate" && format != "array") runtime_fail("unknown format: " + format);
  if (field != "real" && field != "integer" && field != "complex" && field != "pattern")
    runtime_fail("unknown field: " + field);
  if (symmetry != "general" && symmetry != "symmetric" &&
      symmetry != "skew-symmetric" && symmetry != "hermitian")
    runtime_fail("unknown symmetry: " + symmetry);
  // read_matrix_market (matrix_market.hpp:59-71)
  if (format != "coordinate") runtime_fail("only coordinate format is supported");
  if (field == "complex") runtime_fail("complex field is not yet supported");
  *symmetric = symmetry == "symmetric";
  *pattern = field == "pattern";
}

// `is >> integer` from position *s: leading blanks skipped; 0 and no
// advance when nothing parses (the stream's failed extraction)
int64_t take_int(const char** s) {
  const char* p = *s;
  while (*p == ' ' || *p == '\t' || *p == '\r') ++p;
  char* e = nullptr;
  errno = 0;
  const long long v = std::strtoll(p, &e, 10);
  if (e == p) return 0;
  *s = e;
  return (int64_t)v;
}

// `is >> double`: decimal and exponent forms (not inf/nan, which the
// stream does not read)
double take_double(const char** s) {
  const char* p = *s;
  while (*p == ' ' || *p == '\t' || *p == '\r') ++p;
  const char c0 = *p, c1 = c0 ? p[1] : 0;
  const bool numeric = std::isdigit((unsigned char)c0) || c0 == '.' ||
                       ((c0 == '+' || c0 == '-') &&
                        (std::isdigit((unsigned char)c1) || c1 == '.'));
  if (!numeric) return 0.0;
  char* e = nullptr;
  const double v = std::strtod(p, &e);
  if (e == p) return 0.0;
  *s = e;
  return v;
}

spk_csr_s* read_text(const char* text, size_t len) {
  LineReader in{text, text + len};
  std::string line;
  if (!in.next(line)) runtime_fail("unexpected end of input");
  bool symmetric = false, pattern = false;
  parse_banner(line, &symmetric, &pattern);
  while (in.next(line))
    if (!line.empty() && line[0] != '%') break;
  const char* s = line.c_str();
  const int64_t rows = take_int(&s), cols = take_int(&s), nnz = take_int(&s);
  SPK_REQUIRE(rows > 1 && cols > 1, SPK_LOGIC_ERROR, "invalid matrix shape");
  SPK_REQUIRE(nnz >= 0 && nnz < INT32_MAX / 2, SPK_INVALID_ARGUMENT,
              "entry count out of range");
  std::vector<int64_t> r, c;
  std::vector<double> v;
  r.reserve((size_t)(symmetric ? 2 * nnz : nnz));
  c.reserve(r.capacity());
  v.reserve(r.capacity());
  for (int64_t k = 0; k < nnz; ++k) {
    in.next(line);  // past the end: an empty line, index 0 - 1 (Index.cpp:8)
    const char* q = line.c_str();
    const int64_t i = take_int(&q) - 1, j = take_int(&q) - 1;
    const double x = pattern ? 1.0 : take_double(&q);
    r.push_back(i);
    c.push_back(j);
    v.push_back(x);
    if (symmetric && i != j) {
      r.push_back(j);
      c.push_back(i);
      v.push_back(x);
    }
  }
  const int64_t m = (int64_t)r.size();
  DevBuf<int64_t> dr((size_t)(m > 0 ? m : 1)), dc((size_t)(m > 0 ? m : 1));
  DevBuf<double> dv((size_t)(m > 0 ? m : 1));
  hipStream_t st = stream();
  if (m > 0) {
    SPK_HIP_TRY(hipMemcpyAsync(dr.p, r.data(), m * 8, hipMemcpyHostToDevice, st));
    SPK_HIP_TRY(hipMemcpyAsync(dc.p, c.data(), m * 8, hipMemcpyHostToDevice, st));
    SPK_HIP_TRY(hipMemcpyAsync(dv.p, v.data(), m * 8, hipMemcpyHostToDevice, st));
  }
  spk_csr out = nullptr;
  const spk_status rc = spk_csr_assemble(rows, cols, m, dr.p, dc.p, dv.p, &out);
  if (rc != SPK_OK) throw Error{rc, spk_last_error()};
  return out;
}

// write_matrix_market (matrix_market.hpp:123-148)
std::string write_text(spk_csr_s& A) {
  const int64_t n = A.nrows, nnz = A.nnz;
  std::vector<int32_t> ci((size_t)(nnz > 0 ? nnz : 1));
  std::vector<double> v((size_t)(nnz > 0 ? nnz : 1));
  hipStream_t st = stream();
  if (nnz > 0) {
    SPK_HIP_TRY(hipMemcpyAsync(ci.data(), A.ci.p, nnz * 4, hipMemcpyDeviceToHost, st));
    SPK_HIP_TRY(hipMemcpyAsync(v.data(), A.v64.p, nnz * 8, hipMemcpyDeviceToHost, st));
  }
  SPK_HIP_TRY(hipStreamSynchronize(st));
  std::string out = "%%MatrixMarket matrix coordinate real general\n";
  char buf[96];
  std::snprintf(buf, sizeof buf, "%lld %lld %lld\n", (long long)A.nrows,
                (long long)A.ncols, (long long)nnz);
  out += buf;
  out.reserve(out.size() + (size_t)nnz * 24);
  for (int64_t i = 0; i < n; ++i)
    for (int32_t k = A.h_rp[(size_t)i]; k < A.h_rp[(size_t)i + 1]; ++k) {
      std::snprintf(buf, sizeof buf, "%lld %d %g\n", (long long)(i + 1),
                    ci[(size_t)k] + 1, v[(size_t)k]);
      out += buf;
    }
  return out;
}

}  // namespace
}  // namespace spk

using namespace spk;

extern "C" spk_status spk_csr_read_matrix_market_text(const char* text, size_t len,
                                                      spk_csr* out) {
  return guarded([&] {
    SPK_REQUIRE(out && (text || len == 0), SPK_INVALID_ARGUMENT, "null argument");
    *out = read_text(text ? text : "", len);
  });
}

extern "C" spk_status spk_csr_read_matrix_market(const char* path, spk_csr* out) {
  return guarded([&] {
    SPK_REQUIRE(out && path, SPK_INVALID_ARGUMENT, "null argument");
    std::ifstream f(path, std::ios::binary);
    if (!f) runtime_fail(std::string("cannot open file: ") + path);
    std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    *out = read_text(text.data(), text.size());
  });
}

extern "C" spk_status spk_csr_write_matrix_market_text(spk_csr A, char* buf,
                                                       size_t cap, size_t* len) {
  return guarded([&] {
    SPK_REQUIRE(A && len, SPK_INVALID_ARGUMENT, "null argument");
    const std::string t = write_text(*A);
    *len = t.size();
    if (buf && cap > t.size()) std::memcpy(buf, t.c_str(), t.size() + 1);
  });
}

extern "C" spk_status spk_csr_write_matrix_market(spk_csr A, const char* path) {
  return guarded([&] {
    SPK_REQUIRE(A && path, SPK_INVALID_ARGUMENT, "null argument");
    std::ofstream f(path, std::ios::binary);
    if (!f) runtime_fail(std::string("cannot open file: ") + path);
    const std::string t = write_text(*A);
    f.write(t.data(), (std::streamsize)t.size());
    if (!f) runtime_fail(std::string("write failed: ") + path);
  });
}
