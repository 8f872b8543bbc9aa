// Reads a catalog with the PCF command line's reader (desi_amd/cli/pcf_io.h)
// using a given number of threads and writes the n x 4 FP32 values raw.  Used
// by test_reader_cpu.py.
//   reader_check <catalog> <n> <threads> <out.f32>
// Exit 0 and the largest coordinate on stdout, or 1 and the reader's message.
#include <cstdio>
#include <iostream>
#include <vector>

#include "../desi_amd/cli/pcf_io.h"

int main(int argc, char** argv) {
  if (argc != 5) return 2;
  std::vector<float> v;
  double mx = 0;
  std::string err;
  if (!pcfcli::read_catalog(argv[1], std::atol(argv[2]), &v, &mx, &err, std::atoi(argv[3]))) {
    std::cout << err << std::endl;
    return 1;
  }
  FILE* f = std::fopen(argv[4], "wb");
  if (!f) return 3;
  std::fwrite(v.data(), sizeof(float), v.size(), f);
  std::fclose(f);
  std::printf("%.17g\n", mx);
  return 0;
}
