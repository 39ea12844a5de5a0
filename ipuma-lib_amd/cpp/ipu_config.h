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

}  // namespace ipu
#endif
