// fasta.cpp -- FASTA records and dataset files, as the reference reads them
// (src/swatlib/fasta.cpp:15-51, ipusw/run_comparison.hpp:22-61).
#include "fasta.h"

#include <algorithm>
#include <cctype>
#include <stdexcept>

namespace {
void right_trim(std::string &s) {
  while (!s.empty() && std::isspace((unsigned char)s.back())) s.pop_back();
}
void drop_space(std::string &s) {
  s.erase(std::remove_if(s.begin(), s.end(), [](unsigned char c) { return std::isspace(c); }),
          s.end());
}
}  // namespace

std::istream &operator>>(std::istream &input, swatlib::Fasta &f) {
  std::string line;
  bool header = false;
  f.sequence.clear();
  std::streampos before = input.tellg();
  while (std::getline(input, line)) {
    const bool is_header = !line.empty() && line[0] == '>';
    if (is_header && header) {  // the next record: give its header back
      input.seekg(before, std::ios_base::beg);
      break;
    }
    if (is_header) {
      header = true;
      right_trim(line);
      f.description = line;
    } else if (header) {
      drop_space(line);
      f.sequence += line;
    }
    before = input.tellg();
  }
  return input;
}

swatlib::Fasta swatlib::readFasta(std::ifstream &rf) {
  Fasta f;
  rf >> f;
  return f;
}

swatlib::Fasta swatlib::readFasta(const std::string &path) {
  std::ifstream rf(path);
  return readFasta(rf);
}

void ipu::loadSequences(const std::string &path, std::vector<std::string> &sequences) {
  std::ifstream in(path);
  if (!in.good()) throw std::runtime_error("File not found: " + path);
  std::string line;
  while (std::getline(in, line)) sequences.push_back(line);
}

void ipu::load_data(const std::string &path, std::vector<std::string> &seqs, int count) {
  if (path.size() >= 4 && path.compare(path.size() - 4, 4, ".txt") == 0) {
    loadSequences(path, seqs);
    return;
  }
  std::ifstream is(path);
  if (is.fail()) throw std::runtime_error("Opening file at " + path + " failed.");
  swatlib::Fasta f;
  for (int i = 0; !count || i < count; ++i) {
    if (is >> f) {
      seqs.push_back(f.sequence);
    } else {
      seqs.push_back(f.sequence);
      break;
    }
  }
}
