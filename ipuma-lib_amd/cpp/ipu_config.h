// ipu_config.h -- ipu::SWConfig and ipu::IPUAlgoConfig with the reference's
// fields, defaults and buffer-size helpers (src/ipu_config.h:11-42,
// src/ipu_config.cpp:6-14).  On MI355X the tile fields are capacity hints.
#ifndef IPUMA_AMD_IPU_CONFIG_H
#define IPUMA_AMD_IPU_CONFIG_H
#include <string>

#include "types.h"

namespace ipu {

struct SWConfig {
  int gapInit = 0;
  int gapExtend = -1;
  int matchValue = 1;
  int mismatchValue = -1;
  int ambiguityValue = -1;
  swatlib::Similarity similarity = swatlib::Similarity::nucleicAcid;
  swatlib::DataType datatype = swatlib::DataType::nucleicAcid;
};

struct IPUAlgoConfig {
  int tilesUsed = 1;
  int maxAB = 300;
  int maxBatches = 20;
  int bufsize = 3000;
  VertexType vtype = VertexType::cpp;
  Algorithm fillAlgo = Algorithm::fillFirst;
  bool forwardOnly = false;
  bool useRemoteBuffer = false;
  int transmissionPrograms = 1;
  int ioTiles = 0;

  int getTotalNumberOfComparisons() const;
  int getBufsize32b() const;
  int getTotalBufsize32b() const;
  int getMetaBufferSize32b() const;
  int getLenBufferSize32b() const;
  int getInputBufferSize32b() const;
};

// JSON text with the reference's keys (src/ipu_config.cpp:16-80).  The
// reference writes mismatchValue from ambiguityValue (:21); this writes the
// mismatch value.
std::string to_json(const SWConfig &c);
std::string to_json(const IPUAlgoConfig &c);
// Parse the same keys (flat object); throws std::runtime_error on a missing key.
SWConfig sw_config_from_json(const std::string &text);
IPUAlgoConfig algo_config_from_json(const std::string &text);

// The reference's JSON conversions for a JSON library that finds them by
// argument-dependent lookup (nlohmann::json: j = cfg; j.get<SWConfig>()), as
// src/ipu_config.h:47-59 declares them for nlohmann::json; templates on the
// JSON type, so no JSON header is needed here.  Same keys; the enums are
// their config strings ("na", "fillfirst", "multiasm", ...).  to_json writes
// the mismatch value under "mismatchValue" (the reference writes the
// ambiguity value there, src/ipu_config.cpp:21).
template <class Json>
void to_json(Json &j, const Algorithm &a) {
  j = algorithmToConfigString(a);
}
template <class Json>
void from_json(const Json &j, Algorithm &a) {
  a = strToAlgorithm(j.template get<std::string>());
}
template <class Json>
void to_json(Json &j, const VertexType &t) {
  j = vertexTypeToConfigString(t);
}
template <class Json>
void from_json(const Json &j, VertexType &t) {
  t = strToVertexType(j.template get<std::string>());
}
template <class Json>
void to_json(Json &j, const SWConfig &c) {
  j = Json::object();
  j["gapInit"] = c.gapInit;
  j["gapExtend"] = c.gapExtend;
  j["matchValue"] = c.matchValue;
  j["mismatchValue"] = c.mismatchValue;
  j["ambiguityValue"] = c.ambiguityValue;
  j["similarity"] = swatlib::similarityToStr(c.similarity);
  j["datatype"] = swatlib::dataTypeToStr(c.datatype);
}
template <class Json>
void from_json(const Json &j, SWConfig &c) {
  c.gapInit = j.at("gapInit").template get<int>();
  c.gapExtend = j.at("gapExtend").template get<int>();
  c.matchValue = j.at("matchValue").template get<int>();
  c.mismatchValue = j.at("mismatchValue").template get<int>();
  c.ambiguityValue = j.at("ambiguityValue").template get<int>();
  c.similarity = swatlib::strToSimilarity(j.at("similarity").template get<std::string>());
  c.datatype = swatlib::strToDataType(j.at("datatype").template get<std::string>());
}
template <class Json>
void to_json(Json &j, const IPUAlgoConfig &c) {
  j = Json::object();
  j["tilesUsed"] = c.tilesUsed;
  j["maxAB"] = c.maxAB;
  j["maxBatches"] = c.maxBatches;
  j["bufsize"] = c.bufsize;
  j["vtype"] = vertexTypeToConfigString(c.vtype);
  j["fillAlgo"] = algorithmToConfigString(c.fillAlgo);
  j["forwardOnly"] = c.forwardOnly;
  j["useRemoteBuffer"] = c.useRemoteBuffer;
  j["transmissionPrograms"] = c.transmissionPrograms;
  j["ioTiles"] = c.ioTiles;
}
template <class Json>
void from_json(const Json &j, IPUAlgoConfig &c) {
  c.tilesUsed = j.at("tilesUsed").template get<int>();
  c.maxAB = j.at("maxAB").template get<int>();
  c.maxBatches = j.at("maxBatches").template get<int>();
  c.bufsize = j.at("bufsize").template get<int>();
  c.vtype = strToVertexType(j.at("vtype").template get<std::string>());
  c.fillAlgo = strToAlgorithm(j.at("fillAlgo").template get<std::string>());
  c.forwardOnly = j.at("forwardOnly").template get<bool>();
  c.useRemoteBuffer = j.at("useRemoteBuffer").template get<bool>();
  c.transmissionPrograms = j.at("transmissionPrograms").template get<int>();
  c.ioTiles = j.at("ioTiles").template get<int>();
}

}  // namespace ipu

namespace swatlib {
// src/swatlib/similarity.h:115-116, src/swatlib/encoding.h:18-20
template <class Json>
void to_json(Json &j, const Similarity &s) {
  j = similarityToStr(s);
}
template <class Json>
void from_json(const Json &j, Similarity &s) {
  s = strToSimilarity(j.template get<std::string>());
}
template <class Json>
void to_json(Json &j, const DataType &d) {
  j = dataTypeToStr(d);
}
template <class Json>
void from_json(const Json &j, DataType &d) {
  d = strToDataType(j.template get<std::string>());
}
}  // namespace swatlib
#endif
