// partition.h -- ipu::partition, the reference's bucket maps for the prepared
// input buffer (src/partition.h:9-77): which tile ("bucket") and slot every
// comparison lands in, and where its sequences sit in the tile's byte region.
// Same types, fields and functions, so code that builds or inspects maps
// compiles unchanged; the placement rules are the reference's (fillFirst,
// roundRobin, greedy; src/partition.cpp:105-462), so prepared buffers are
// byte-identical to the reference's for the same inputs.
//
// Differences a caller can observe:
//  * greedy: when the least-loaded bucket is full, the reference copy-assigns
//    the next heap entry over it (`bucket = q.top().get()` through a
//    reference, src/partition.cpp:337, 408), which overwrites that bucket's
//    contents; here the next bucket is taken instead.  Maps where no bucket
//    fills up are identical.
//  * the per-bucket statistics logging (fixed 1,472-entry vectors,
//    src/partition.cpp:271-277, 445-451) is not kept.
#ifndef IPUMA_AMD_PARTITION_H
#define IPUMA_AMD_PARTITION_H
#include <deque>
#include <functional>
#include <queue>
#include <string>
#include <vector>

#include "types.h"

namespace ipu {
namespace partition {

enum class SequenceOrigin { unordered, A, B };
std::string sequenceOriginToString(SequenceOrigin o);

struct ComparisonMapping {
  int comparisonIndex;  // index in the input list
  int sizeA;
  int offsetA;
  int sizeB;
  int offsetB;
  std::string toString();
};

struct SequenceMapping {
  int index;  // index in the input list
  int offset;
  SequenceOrigin origin = SequenceOrigin::unordered;
  std::string toString();
};

struct BucketMapping {
  int bucketIndex;
  int maxLen = 0;      // longest sequence
  int seqSize = 0;     // bytes of sequences held
  int weight[6] = {};  // cells per worker slot (cmps.size() - 1) % 6 (greedy / round robin)
  std::vector<SequenceMapping> seqs;
  std::vector<ComparisonMapping> cmps;
  std::string toString();
};

bool operator<(const BucketMapping &b1, const BucketMapping &b2);
bool operator>(const BucketMapping &b1, const BucketMapping &b2);

struct BucketMap {
  std::vector<BucketMapping> buckets;
  int numBuckets;
  int cmpCapacity;
  int sequenceCapacity;

  BucketMap();
  BucketMap(int nB, int nC, int sC);
  std::string toString();
};

struct BucketData {
  int count;
  int lenSeq;
  int weight;
};

// Min-heap of buckets by their lightest worker slot.
using BucketHeap =
    std::priority_queue<std::reference_wrapper<BucketMapping>,
                        std::deque<std::reference_wrapper<BucketMapping>>,
                        std::greater<std::deque<std::reference_wrapper<BucketMapping>>::value_type>>;

// A/B lists (comparison k = A[k] vs B[k]); the bool forms continue an earlier
// fill (curBucket / heap) and return false when a comparison fits nowhere.
bool fillFirst(BucketMap &map, const RawSequences &A, const RawSequences &B, int indexOffset,
               int &curBucket);
bool roundRobin(BucketMap &map, const RawSequences &A, const RawSequences &B, int indexOffset,
                int &curBucket);
bool greedy(BucketMap &map, const RawSequences &A, const RawSequences &B, int indexOffset,
            BucketHeap &heap);

// Throw std::runtime_error("Out of buckets.") when a comparison fits nowhere.
void fillFirst(BucketMap &map, const RawSequences &A, const RawSequences &B, int indexOffset = 0);
void roundRobin(BucketMap &map, const RawSequences &A, const RawSequences &B, int indexOffset = 0);
void greedy(BucketMap &map, const RawSequences &A, const RawSequences &B, int indexOffset = 0);

// Comparisons over one sequence list; a sequence already in the chosen bucket
// is not stored twice.
void fillFirst(BucketMap &map, const RawSequences &Seqs, const Comparisons &Cmps,
               int indexOffset = 0);
void roundRobin(BucketMap &map, const RawSequences &Seqs, const Comparisons &Cmps,
                int indexOffset = 0);
void greedy(BucketMap &map, const RawSequences &Seqs, const Comparisons &Cmps,
            int indexOffset = 0);

bool fillBuckets(Algorithm algo, BucketMap &map, const RawSequences &A, const RawSequences &B,
                 int indexOffset, int &curBucket, BucketHeap &heap);

}  // namespace partition
}  // namespace ipu
#endif
