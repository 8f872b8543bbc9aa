// pcf_io.h -- catalog reader and histogram writers of the PCF command line.
//
// Catalog format (readme.md:36-40): one point per row, four whitespace
// separated columns x y z w.  The reference reads the first -n rows with
// `ifstream >> float` (cuda/create_grid.cuh:51-53); here the file is read in one
// piece and parsed with strtof, which rounds decimal text to FP32 the same way.
//
// Writers are byte-compatible with the reference's save_histogram1D/2D/3D/5D
// (cuda/create_grid.cuh:141-266): values printed through an ostream with
// setprecision(12) in the default float format; 1-D one value per line; 2-D rows
// of "v " terminated by '\n'; 3-D blocks of such rows, each block followed by
// "\n" and endl; 5-D nested the same way.
#pragma once

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

namespace pcfcli {

// Reads `n` points.  Returns false with a message when the file cannot be
// opened or holds fewer than n complete rows.  `max_component` receives the
// largest of x, y, z over the rows read (starting from 0, as the reference's
// box inference does, cuda/create_grid.cuh:48-72).
inline bool read_catalog(const std::string& path, long n, std::vector<float>* xyzw,
                         double* max_component, std::string* err) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) {
    *err = "Failed to load the file in " + path;
    return false;
  }
  std::string buf;
  char chunk[1 << 16];
  size_t got;
  while ((got = std::fread(chunk, 1, sizeof(chunk), f)) > 0) buf.append(chunk, got);
  std::fclose(f);
  xyzw->assign((size_t)n * 4, 0.f);
  const char* p = buf.c_str();
  double mx = 0;
  for (long i = 0; i < n; ++i) {
    for (int k = 0; k < 4; ++k) {
      char* end;
      errno = 0;
      float v = std::strtof(p, &end);
      if (end == p) {
        *err = path + ": fewer than " + std::to_string(n) + " rows of 4 columns (row " +
               std::to_string(i + 1) + ")";
        return false;
      }
      (*xyzw)[(size_t)i * 4 + k] = v;
      p = end;
    }
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
