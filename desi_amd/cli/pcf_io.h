// pcf_io.h -- catalog reader and histogram writers of the PCF command line.
//
// Catalog format (readme.md:36-40): one point per row, four whitespace
// separated columns x y z w.  The reference reads the first -n rows with
// `ifstream >> float` (cuda/create_grid.cuh:51-53); here the file is read in one
// piece and parsed by several threads with strtof, which rounds decimal text
// to FP32 the same way (read_catalog).
//
// Writers are byte-compatible with the reference's save_histogram1D/2D/3D/5D
// (cuda/create_grid.cuh:141-266): values printed through an ostream with
// setprecision(12) in the default float format; 1-D one value per line; 2-D rows
// of "v " terminated by '\n'; 3-D blocks of such rows, each block followed by
// "\n" and endl; 5-D nested the same way.
#pragma once

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <iomanip>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

namespace pcfcli {

namespace detail {

inline bool is_space(char c) {
  return c == ' ' || c == '\n' || c == '\t' || c == '\r' || c == '\v' || c == '\f';
}

// Parses the whitespace-separated tokens of buf[b, e) in order into out[k0 ...],
// stopping at out index `lim`.  A token must be one whole float (strtof: the
// same correctly rounded decimal -> FP32 conversion as `ifstream >> float`).
// Returns the index of the first bad token, or -1.
inline long parse_tokens(const char* buf, size_t b, size_t e, long k0, long lim, float* out) {
  size_t p = b;
  long k = k0;
  char tok[64];
  std::string big;
  while (k < lim) {
    while (p < e && is_space(buf[p])) ++p;
    if (p >= e) break;
    size_t q = p;
    while (q < e && !is_space(buf[q])) ++q;
    const size_t len = q - p;
    char* t = tok;
    if (len >= sizeof(tok)) {
      big.assign(buf + p, len);
      t = &big[0];
    } else {
      std::memcpy(tok, buf + p, len);
      tok[len] = 0;
    }
    char* end;
    const float v = std::strtof(t, &end);
    if (end != t + len) return k;
    out[k++] = v;
    p = q;
  }
  return -1;
}

inline long count_tokens(const char* buf, size_t b, size_t e) {
  long n = 0;
  bool in = false;
  for (size_t p = b; p < e; ++p) {
    const bool sp = is_space(buf[p]);
    if (!sp && !in) ++n;
    in = !sp;
  }
  return n;
}

}  // namespace detail

// Reads `n` points.  Returns false with a message when the file cannot be
// opened or holds fewer than n complete rows.  `max_component` receives the
// largest of x, y, z over the rows read (starting from 0, as the reference's
// box inference does, cuda/create_grid.cuh:48-72).
//
// Text catalogs (the reference's format) are parsed in parallel: the file is
// cut into slices at whitespace, every thread counts the tokens of its slice,
// a prefix sum places each slice's first token, and the threads convert their
// tokens with strtof.  As with the reference's `ifstream >> float`, line
// breaks carry no meaning: the first 4n tokens are the n points.  A file
// whose name ends in ".bin" holds raw little-endian FP32 rows x, y, z, w (16
// bytes per point), read without conversion.
inline bool read_catalog(const std::string& path, long n, std::vector<float>* xyzw,
                         double* max_component, std::string* err, int threads = 0) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) {
    *err = "Failed to load the file in " + path;
    return false;
  }
  std::string buf;
  std::fseek(f, 0, SEEK_END);
  const long fsz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (fsz > 0) {
    buf.resize((size_t)fsz);
    const size_t got = std::fread(&buf[0], 1, (size_t)fsz, f);
    buf.resize(got);
  }
  std::fclose(f);
  xyzw->assign((size_t)n * 4, 0.f);
  const bool binary = path.size() >= 4 && path.compare(path.size() - 4, 4, ".bin") == 0;
  if (binary) {
    if (buf.size() < (size_t)n * 16) {
      *err = path + ": fewer than " + std::to_string(n) + " points of 16 bytes";
      return false;
    }
    if (n > 0) std::memcpy(xyzw->data(), buf.data(), (size_t)n * 16);
  } else {
    const size_t sz = buf.size();
    const char* b = buf.data();
    int T = threads > 0 ? threads : (int)std::thread::hardware_concurrency();
    T = std::max(1, std::min({T, 16, (int)(sz / (1 << 20)) + 1}));
    // Slice boundaries moved forward to the next whitespace character.
    std::vector<size_t> cut(T + 1, sz);
    cut[0] = 0;
    for (int t = 1; t < T; ++t) {
      size_t p = std::max(cut[t - 1], sz * (size_t)t / (size_t)T);
      while (p < sz && !detail::is_space(b[p])) ++p;
      cut[t] = p;
    }
    std::vector<long> cnt(T, 0), first(T + 1, 0), bad(T, -1);
    auto pool = [&](auto fn) {
      std::vector<std::thread> th;
      for (int t = 1; t < T; ++t) th.emplace_back(fn, t);
      fn(0);
      for (auto& x : th) x.join();
    };
    pool([&](int t) { cnt[t] = detail::count_tokens(b, cut[t], cut[t + 1]); });
    for (int t = 0; t < T; ++t) first[t + 1] = first[t] + cnt[t];
    const long need = 4 * n;
    if (first[T] < need) {
      *err = path + ": fewer than " + std::to_string(n) + " rows of 4 columns (row " +
             std::to_string(first[T] / 4 + 1) + ")";
      return false;
    }
    float* out = xyzw->data();
    pool([&](int t) {
      if (first[t] < need) bad[t] = detail::parse_tokens(b, cut[t], cut[t + 1], first[t], need, out);
    });
    for (int t = 0; t < T; ++t)
      if (bad[t] >= 0) {
        *err = path + ": fewer than " + std::to_string(n) + " rows of 4 columns (row " +
               std::to_string(bad[t] / 4 + 1) + ")";
        return false;
      }
  }
  double mx = 0;
  for (long i = 0; i < n; ++i) {
    const float* q = &(*xyzw)[(size_t)i * 4];
    double m = q[0] > q[1] ? (q[0] > q[2] ? q[0] : q[2]) : (q[1] > q[2] ? q[1] : q[2]);
    if (m > mx) mx = m;
  }
  *max_component = mx;
  return true;
}

inline bool open_out(const std::string& path, std::ofstream& f) {
  f.open(path.c_str(), std::ios::out | std::ios::binary);
  if (f.fail()) {
    std::cout << "Failed to save the the histogram in " << path << std::endl;
    return false;
  }
  return true;
}

inline bool save_histogram1D(const std::string& path, int bns, const double* h) {
  std::ofstream f;
  if (!open_out(path, f)) return false;
  for (int i = 0; i < bns; i++) f << std::setprecision(12) << h[i] << std::endl;
  return true;
}

inline bool save_histogram2D(const std::string& path, int bns, const double* h) {
  std::ofstream f;
  if (!open_out(path, f)) return false;
  for (int i = 0; i < bns; i++) {
    for (int j = 0; j < bns; j++) f << std::setprecision(12) << h[i * bns + j] << ' ';
    f << "\n";
  }
  return true;
}

inline bool save_histogram3D(const std::string& path, int bns, const double* h) {
  std::ofstream f;
  if (!open_out(path, f)) return false;
  for (int i = 0; i < bns; i++) {
    for (int j = 0; j < bns; j++) {
      for (int k = 0; k < bns; k++)
        f << std::setprecision(12) << h[((long)i * bns + j) * bns + k] << ' ';
      f << "\n";
    }
    f << "\n" << std::endl;
  }
  return true;
}

inline bool save_histogram5D(const std::string& path, int bns, const double* h) {
  std::ofstream f;
  if (!open_out(path, f)) return false;
  long b = bns;
  for (long i = 0; i < b; i++) {
    for (long j = 0; j < b; j++) {
      for (long k = 0; k < b; k++) {
        for (long l = 0; l < b; l++) {
          for (long m = 0; m < b; m++)
            f << std::setprecision(12) << h[(((i * b + j) * b + k) * b + l) * b + m] << ' ';
          f << "\n";
        }
        f << "\n";
      }
      f << "\n";
    }
    f << "\n" << std::endl;
  }
  return true;
}

}  // namespace pcfcli
