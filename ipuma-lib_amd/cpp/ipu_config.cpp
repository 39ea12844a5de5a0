// ipu_config.cpp -- buffer-size helpers and flat JSON of the configs.
#include "ipu_config.h"

#include <cmath>
#include <cstdlib>
#include <sstream>
#include <stdexcept>

namespace ipu {

int IPUAlgoConfig::getBufsize32b() const { return (bufsize + 3) / 4; }
int IPUAlgoConfig::getTotalNumberOfComparisons() const { return tilesUsed * maxBatches; }
int IPUAlgoConfig::getMetaBufferSize32b() const { return getTotalNumberOfComparisons() * 4; }
int IPUAlgoConfig::getLenBufferSize32b() const { return getTotalNumberOfComparisons(); }
int IPUAlgoConfig::getTotalBufsize32b() const { return tilesUsed * getBufsize32b(); }
int IPUAlgoConfig::getInputBufferSize32b() const {
  return getTotalBufsize32b() + getMetaBufferSize32b();
}

std::string to_json(const SWConfig &c) {
  std::ostringstream o;
  o << "{\"gapInit\":" << c.gapInit << ",\"gapExtend\":" << c.gapExtend
    << ",\"matchValue\":" << c.matchValue << ",\"mismatchValue\":" << c.mismatchValue
    << ",\"ambiguityValue\":" << c.ambiguityValue << ",\"similarity\":\""
    << swatlib::similarityToStr(c.similarity) << "\",\"datatype\":\""
    << swatlib::dataTypeToStr(c.datatype) << "\"}";
  return o.str();
}

std::string to_json(const IPUAlgoConfig &c) {
  std::ostringstream o;
  o << "{\"tilesUsed\":" << c.tilesUsed << ",\"maxAB\":" << c.maxAB
    << ",\"maxBatches\":" << c.maxBatches << ",\"bufsize\":" << c.bufsize << ",\"vtype\":\""
    << vertexTypeToConfigString(c.vtype) << "\",\"fillAlgo\":\""
    << algorithmToConfigString(c.fillAlgo) << "\",\"forwardOnly\":"
    << (c.forwardOnly ? "true" : "false") << ",\"useRemoteBuffer\":"
    << (c.useRemoteBuffer ? "true" : "false") << ",\"transmissionPrograms\":"
    << c.transmissionPrograms << ",\"ioTiles\":" << c.ioTiles << "}";
  return o.str();
}

namespace {
// Value text of "key" in a flat JSON object (strings without escapes).
std::string field(const std::string &text, const char *key) {
  const std::string k = std::string("\"") + key + "\"";
  size_t p = text.find(k);
  if (p == std::string::npos) throw std::runtime_error(std::string("missing key ") + key);
  p = text.find(':', p + k.size());
  if (p == std::string::npos) throw std::runtime_error(std::string("bad value for ") + key);
  ++p;
  while (p < text.size() && (text[p] == ' ' || text[p] == '\n' || text[p] == '\t')) ++p;
  if (p < text.size() && text[p] == '"') {
    const size_t q = text.find('"', p + 1);
    return text.substr(p + 1, q - p - 1);
  }
  size_t q = p;
  while (q < text.size() && text[q] != ',' && text[q] != '}' && text[q] != '\n') ++q;
  std::string v = text.substr(p, q - p);
  while (!v.empty() && v.back() == ' ') v.pop_back();
  return v;
}
int ival(const std::string &text, const char *key) { return std::atoi(field(text, key).c_str()); }
bool bval(const std::string &text, const char *key) { return field(text, key) == "true"; }
}  // namespace

SWConfig sw_config_from_json(const std::string &text) {
  SWConfig c;
  c.gapInit = ival(text, "gapInit");
  c.gapExtend = ival(text, "gapExtend");
  c.matchValue = ival(text, "matchValue");
  c.mismatchValue = ival(text, "mismatchValue");
  c.ambiguityValue = ival(text, "ambiguityValue");
  c.similarity = swatlib::strToSimilarity(field(text, "similarity"));
  c.datatype = swatlib::strToDataType(field(text, "datatype"));
  return c;
}

IPUAlgoConfig algo_config_from_json(const std::string &text) {
  IPUAlgoConfig c;
  c.tilesUsed = ival(text, "tilesUsed");
  c.maxAB = ival(text, "maxAB");
  c.maxBatches = ival(text, "maxBatches");
  c.bufsize = ival(text, "bufsize");
  c.vtype = strToVertexType(field(text, "vtype"));
  c.fillAlgo = strToAlgorithm(field(text, "fillAlgo"));
  c.forwardOnly = bval(text, "forwardOnly");
  c.useRemoteBuffer = bval(text, "useRemoteBuffer");
  c.transmissionPrograms = ival(text, "transmissionPrograms");
  c.ioTiles = ival(text, "ioTiles");
  return c;
}

}  // namespace ipu
