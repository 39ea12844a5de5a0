// driver.hpp -- the reference's process-global entry for embedding aligners
// (src/driver.hpp:5-67, src/driver.cpp:5-17): compile-time ALN_* scoring and
// KLIGN_* capacity defaults, SW_CONFIGURATION / ALGO_CONFIGURATION, getDriver()
// and init_single_ipu().
#ifndef IPUMA_AMD_DRIVER_HPP
#define IPUMA_AMD_DRIVER_HPP

#ifndef KLIGN_IPU_MAXAB_SIZE
#define KLIGN_IPU_MAXAB_SIZE 200
#endif
#ifndef KLIGN_IPU_TILES
#define KLIGN_IPU_TILES 8832
#endif
#ifndef KLIGN_IPUS_LOCAL
#define KLIGN_IPUS_LOCAL 1
#endif
#ifndef KLIGN_IPU_MAX_BATCHES
#define KLIGN_IPU_MAX_BATCHES 300
#endif
#ifndef KLIGN_IPU_BUFSIZE
#define KLIGN_IPU_BUFSIZE 30000
#endif
#ifndef ALN_GAP_OPENING_COST
#define ALN_GAP_OPENING_COST 1
#endif
#ifndef ALN_GAP_EXTENDING_COST
#define ALN_GAP_EXTENDING_COST 1
#endif
#ifndef ALN_MATCH_SCORE
#define ALN_MATCH_SCORE 1
#endif
#ifndef ALN_MISMATCH_COST
#define ALN_MISMATCH_COST 1
#endif
#ifndef ALN_AMBIGUITY_COST
#define ALN_AMBIGUITY_COST 1
#endif

#include "ipu_batch_affine.h"

// A gap of length k costs open + (k - 1) * extend.
static const ipu::SWConfig SW_CONFIGURATION = {
    -(ALN_GAP_OPENING_COST - ALN_GAP_EXTENDING_COST),
    -ALN_GAP_EXTENDING_COST,
    ALN_MATCH_SCORE,
    -ALN_MISMATCH_COST,
    -ALN_AMBIGUITY_COST,
    swatlib::Similarity::nucleicAcid,
    swatlib::DataType::nucleicAcid,
};

static const ipu::IPUAlgoConfig ALGO_CONFIGURATION = {
    KLIGN_IPU_TILES, KLIGN_IPU_MAXAB_SIZE, KLIGN_IPU_MAX_BATCHES, KLIGN_IPU_BUFSIZE,
    ipu::VertexType::cpp, ipu::Algorithm::fillFirst};

ipu::batchaffine::SWAlgorithm *getDriver();
void init_single_ipu(ipu::SWConfig config, ipu::IPUAlgoConfig algoconfig);

#endif
