// Reads a whitespace-separated histogram and writes it back with the PCF
// command line's writers (desi_amd/cli/pcf_io.h).  Used by test_abi_cpu.py to
// check byte-compatibility with files the reference wrote.
//   writer_check <dim 1|2|3|5> <bins> <in> <out>
#include <fstream>
#include <iostream>
#include <vector>

#include "../desi_amd/cli/pcf_io.h"

int main(int argc, char** argv) {
  if (argc != 5) return 2;
  int dim = std::atoi(argv[1]), bins = std::atoi(argv[2]);
  std::ifstream in(argv[3]);
  std::vector<double> v;
  double x;
  while (in >> x) v.push_back(x);
  bool ok = false;
  if (dim == 1) ok = pcfcli::save_histogram1D(argv[4], bins, v.data());
  if (dim == 2) ok = pcfcli::save_histogram2D(argv[4], bins, v.data());
  if (dim == 3) ok = pcfcli::save_histogram3D(argv[4], bins, v.data());
  if (dim == 5) ok = pcfcli::save_histogram5D(argv[4], bins, v.data());
  return ok ? 0 : 1;
}
