// fasta.h -- swatlib::Fasta and its reader (src/swatlib/fasta.h:9-24), plus the
// dataset loader of the reference's batch driver (ipusw/run_comparison.hpp:
// 22-61): a ".txt" file is one sequence per line, anything else is FASTA.
#ifndef IPUMA_AMD_FASTA_H
#define IPUMA_AMD_FASTA_H
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

namespace swatlib {

struct Fasta {
  std::string description;  // the header line ('>' included), trailing space removed
  std::string sequence;     // the record's lines joined, whitespace removed
};

// The first record of the file / the next record of the stream.
Fasta readFasta(const std::string &path);
Fasta readFasta(std::ifstream &rf);

}  // namespace swatlib

// Reads the next record; lines before the first header are skipped, and the
// stream is left at the next header.
std::istream &operator>>(std::istream &input, swatlib::Fasta &f);

namespace ipu {
// One sequence per line (empty lines give empty sequences); throws
// std::runtime_error when the file cannot be opened (the reference exits).
void loadSequences(const std::string &path, std::vector<std::string> &sequences);
// ".txt": loadSequences; otherwise the FASTA records' sequences, the first
// `count` only when count > 0.  Like the reference, the record being read when
// the file ends is appended, so a file without records yields one empty
// sequence.
void load_data(const std::string &path, std::vector<std::string> &seqs, int count = 0);
}  // namespace ipu
#endif
