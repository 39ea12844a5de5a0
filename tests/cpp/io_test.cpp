// io_test.cpp -- prints what the C++ dataset loader reads from a file:
// "<count>\n" then one "[<sequence>]" line per sequence.  Host only.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "fasta.h"

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  std::vector<std::string> seqs;
  ipu::load_data(argv[1], seqs, argc > 2 ? std::atoi(argv[2]) : 0);
  std::printf("%zu\n", seqs.size());
  for (const auto &s : seqs) std::printf("[%s]\n", s.c_str());
  if (argc > 3) {  // also the first record through readFasta
    const swatlib::Fasta f = swatlib::readFasta(std::string(argv[1]));
    std::printf("{%s}{%s}\n", f.description.c_str(), f.sequence.c_str());
  }
  return 0;
}
