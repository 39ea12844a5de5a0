// types.h -- ipu::RawSequences / Comparison / VertexType / Algorithm, as the
// reference declares them (src/types.h:9-29) so existing callers compile.
#ifndef IPUMA_AMD_TYPES_H
#define IPUMA_AMD_TYPES_H
#include <cstdint>
#include <string>
#include <vector>

#define IPU_JSON_LOG_TAG "IPUSWLOG"

namespace swatlib {
// src/swatlib/similarity.h:108-113 and src/swatlib/encoding.h:12-16
enum class Similarity { nucleicAcid, blosum50, blosum62, simple };
enum class DataType { nucleicAcid, aminoAcid, string };
Similarity strToSimilarity(const std::string &s);
std::string similarityToStr(Similarity s);
DataType strToDataType(const std::string &s);
std::string dataTypeToStr(DataType d);
}  // namespace swatlib

namespace ipu {
typedef std::vector<std::string> RawSequences;
typedef std::vector<uint8_t> EncSequences;

struct __attribute__((__packed__)) Comparison {
  int32_t indexA;
  int32_t indexB;
};
typedef std::vector<Comparison> Comparisons;

enum class VertexType { cpp, assembly, multi, multiasm };
static const std::vector<std::string> vertexTypeNames = {"cpp", "assembly", "multi", "multiasm"};
enum class Algorithm { fillFirst, roundRobin, greedy };
static const std::vector<std::string> algoNames = {"fillfirst", "roundrobin", "greedy"};
static const std::string typeLabels[] = {"SWAffine", "SWAffineAsm", "MultiSWAffine",
                                         "MultiSWAffineAsm"};

std::string vertexTypeToIpuLabel(VertexType v);
std::string vertexTypeToConfigString(VertexType v);
std::string algorithmToConfigString(Algorithm a);
VertexType strToVertexType(std::string s);
Algorithm strToAlgorithm(std::string s);
}  // namespace ipu
#endif
