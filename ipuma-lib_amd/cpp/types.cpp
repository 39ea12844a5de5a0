// types.cpp -- string maps of the enums (behaviour of src/types.cpp:4-19 and
// the swatlib string converters, src/swatlib/similarity.cpp:45-72,
// src/swatlib/encoding.cpp:4-29).
#include "types.h"

#include <stdexcept>

namespace swatlib {
Similarity strToSimilarity(const std::string &s) {
  if (s == "na") return Similarity::nucleicAcid;
  if (s == "blosum50") return Similarity::blosum50;
  if (s == "blosum62") return Similarity::blosum62;
  if (s == "simple") return Similarity::simple;
  throw std::runtime_error("Unknown similarity matrix type.");
}
std::string similarityToStr(Similarity s) {
  switch (s) {
    case Similarity::nucleicAcid: return "na";
    case Similarity::blosum50: return "blosum50";
    case Similarity::blosum62: return "blosum62";
    case Similarity::simple: return "simple";
  }
  throw std::runtime_error("Invalid similarity type.");
}
DataType strToDataType(const std::string &s) {
  if (s == "na") return DataType::nucleicAcid;
  if (s == "aa") return DataType::aminoAcid;
  if (s == "str") return DataType::string;
  throw std::runtime_error("Invalid dataType: " + s);
}
std::string dataTypeToStr(DataType d) {
  switch (d) {
    case DataType::nucleicAcid: return "na";
    case DataType::aminoAcid: return "aa";
    case DataType::string: return "str";
  }
  throw std::runtime_error("Invalid DataType");
}
}  // namespace swatlib

namespace ipu {
std::string vertexTypeToIpuLabel(VertexType v) { return typeLabels[static_cast<int>(v)]; }
std::string vertexTypeToConfigString(VertexType v) { return vertexTypeNames[static_cast<int>(v)]; }
std::string algorithmToConfigString(Algorithm a) { return algoNames[static_cast<int>(a)]; }
VertexType strToVertexType(std::string s) {
  for (size_t i = 0; i < vertexTypeNames.size(); ++i)
    if (vertexTypeNames[i] == s) return static_cast<VertexType>(i);
  throw std::runtime_error("Invalid vertex type string: " + s);
}
Algorithm strToAlgorithm(std::string s) {
  for (size_t i = 0; i < algoNames.size(); ++i)
    if (algoNames[i] == s) return static_cast<Algorithm>(i);
  throw std::runtime_error("Invalid algorithm string: " + s);
}
}  // namespace ipu
