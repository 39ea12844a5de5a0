// partition.cpp -- bucket maps of the prepared input buffer: the reference's
// placement rules (src/partition.cpp:105-462), restated.
//
// Rules kept exactly, since they fix the bytes of a prepared buffer:
//  * fillFirst: each comparison goes to the first bucket, from the current
//    one on (wrapping), with room for one more comparison and its sequences;
//    the current bucket moves there.
//  * roundRobin: the same search, then the current bucket advances by one.
//  * greedy: comparisons by |A|*|B| descending (std::sort, as the reference),
//    each into the bucket whose lightest worker slot is lightest (a
//    std::priority_queue min-heap); full buckets are set aside and pushed back
//    afterwards.  Worker slot k = (comparisons in the bucket - 1) % 6 collects
//    |A|*|B| (int arithmetic, as the reference).
//  * in the comparison-list forms a sequence already stored in the chosen
//    bucket is referenced, not stored again.
#include "partition.h"

#include <algorithm>
#include <cstdint>
#include <sstream>
#include <stdexcept>
#include <utility>

namespace ipu {
namespace partition {
namespace {

int min_weight(const int w[6]) {
  int m = w[0];
  for (int i = 1; i < 6; ++i) m = std::min(m, w[i]);
  return m;
}

// int += size_t product, wrapping like the reference's int accumulation
void add_weight(BucketMapping &b, size_t cells) {
  int &w = b.weight[(b.cmps.size() - 1) % 6];
  w = (int)(uint32_t)((uint64_t)(uint32_t)w + (uint64_t)cells);
}

bool fits(const BucketMapping &b, const BucketMap &map, size_t bytes) {
  return b.cmps.size() + 1 <= (size_t)map.cmpCapacity &&
         (int64_t)b.seqSize + (int64_t)bytes <= (int64_t)map.sequenceCapacity;
}

// Places comparison i (A/B lists) into bucket b: A then B appended.
void place_ab(BucketMapping &b, int ci, size_t aLen, size_t bLen) {
  b.maxLen = std::max(b.maxLen, (int)std::max(aLen, bLen));
  const int offA = b.seqSize, offB = b.seqSize + (int)aLen;
  b.seqs.push_back({ci, offA, SequenceOrigin::A});
  b.seqs.push_back({ci, offB, SequenceOrigin::B});
  b.cmps.push_back({ci, (int)aLen, offA, (int)bLen, offB});
  b.seqSize += (int)(aLen + bLen);
}

// First bucket from `cur` on (wrapping) with room, or -1.
int search_from(const BucketMap &map, int cur, size_t bytes) {
  for (int off = 0; off < map.numBuckets; ++off) {
    const int w = (cur + off) % map.numBuckets;
    if (fits(map.buckets[w], map, bytes)) return w;
  }
  return -1;
}

// Comparison-list placement state: the bytes a comparison adds to a bucket
// and the offsets of sequences the bucket already holds.
struct Need {
  int aLen, bLen;
  int effA = 0, effB = 0;  // bytes to add (0: already present)
  int offA = 0, offB = 0;  // offsets when present
};

void lookup(const BucketMapping &b, int keyA, int keyB, Need &nd) {
  nd.effA = nd.aLen;
  nd.effB = nd.bLen;
  for (const auto &s : b.seqs) {  // the last match wins, as in the reference
    if (s.index == keyA) {
      nd.effA = 0;
      nd.offA = s.offset;
    }
    if (s.index == keyB) {
      nd.effB = 0;
      nd.offB = s.offset;
    }
  }
}

// Next bucket from `cur` on with room for the comparison (its present
// sequences do not count), src/partition.cpp:105-138.
int next_bucket_mn(BucketMap &map, int cur, const Comparison &cmp, Need &nd, int indexOffset) {
  for (int off = 0; off < map.numBuckets; ++off) {
    const int w = (cur + off) % map.numBuckets;
    const BucketMapping &b = map.buckets[w];
    lookup(b, indexOffset + cmp.indexA, indexOffset + cmp.indexB, nd);
    if (b.cmps.size() + 1 <= (size_t)map.cmpCapacity &&
        (int64_t)b.seqSize + nd.effA + nd.effB <= (int64_t)map.sequenceCapacity)
      return w;
  }
  throw std::runtime_error("Out of buckets.");
}

void place_mn(BucketMapping &b, Need &nd, const Comparison &cmp, int ci, int indexOffset) {
  b.maxLen = std::max(b.maxLen, std::max(nd.aLen, nd.bLen));
  if (nd.effA) {
    nd.offA = b.seqSize;
    b.seqs.push_back({indexOffset + cmp.indexA, nd.offA, SequenceOrigin::unordered});
  }
  if (nd.effB) {
    nd.offB = b.seqSize + nd.effA;
    b.seqs.push_back({indexOffset + cmp.indexB, nd.offB, SequenceOrigin::unordered});
  }
  b.cmps.push_back({ci, nd.aLen, nd.offA, nd.bLen, nd.offB});
  b.seqSize += nd.effA + nd.effB;
}

// Comparison order of greedy: |A|*|B| descending (the reference's pair<int,
// int> list and std::sort, so equal products keep its order).
std::vector<std::pair<int, int>> by_cells(size_t n, const std::function<size_t(size_t)> &cells) {
  std::vector<std::pair<int, int>> srts(n);
  for (size_t i = 0; i < n; ++i) srts[i] = {(int)cells(i), (int)i};
  std::sort(srts.begin(), srts.end(),
            [](std::pair<int, int> &a, std::pair<int, int> &b) { return a.first > b.first; });
  return srts;
}

// Takes the lightest bucket that passes `ok`; full ones go back afterwards.
template <class OK>
BucketMapping *take_lightest(BucketHeap &q, const OK &ok) {
  std::deque<std::reference_wrapper<BucketMapping>> full;
  BucketMapping *chosen = nullptr;
  while (!q.empty()) {
    BucketMapping &b = q.top().get();
    q.pop();
    if (ok(b)) {
      chosen = &b;
      break;
    }
    full.push_back(std::ref(b));
  }
  for (auto b : full) q.push(b);
  return chosen;
}

}  // namespace

std::string sequenceOriginToString(SequenceOrigin o) {
  switch (o) {
    case SequenceOrigin::A: return "SequenceOrigin::A";
    case SequenceOrigin::B: return "SequenceOrigin::B";
    case SequenceOrigin::unordered: return "SequenceOrigin::unordered";
  }
  throw std::runtime_error("Unhandled sequence type.");
}

std::string ComparisonMapping::toString() {
  std::stringstream ss;
  ss << "CM[" << comparisonIndex << ": a(l" << sizeA << " o" << offsetA << ") b(l" << sizeB
     << " o" << offsetB << ")]";
  return ss.str();
}

std::string SequenceMapping::toString() {
  std::stringstream ss;
  ss << "SM[" << index << ": o" << offset << " t" << sequenceOriginToString(origin) << "]";
  return ss.str();
}

std::string BucketMapping::toString() {
  std::stringstream ss;
  ss << "BMapping[" << bucketIndex << ": cmps(" << cmps.size() << ") maxLen(" << maxLen
     << ") seqSize(" << seqSize << ") weight(";
  for (int i = 0; i < 6; ++i) ss << (i ? "," : "") << weight[i];
  ss << ")]";
  return ss.str();
}

std::string BucketMap::toString() {
  std::stringstream ss;
  ss << "BMap[" << numBuckets << ": maxCmps(" << cmpCapacity << ") bufsize(" << sequenceCapacity
     << ")]";
  return ss.str();
}

BucketMap::BucketMap() : numBuckets(0), cmpCapacity(0), sequenceCapacity(0) {}

BucketMap::BucketMap(int nB, int cC, int sC) : numBuckets(nB), cmpCapacity(cC), sequenceCapacity(sC) {
  buckets.resize(nB > 0 ? nB : 0);
  for (int i = 0; i < nB; ++i) buckets[i].bucketIndex = i;
}

bool operator>(const BucketMapping &b1, const BucketMapping &b2) {
  return min_weight(b1.weight) > min_weight(b2.weight);
}
bool operator<(const BucketMapping &b1, const BucketMapping &b2) {
  return min_weight(b1.weight) < min_weight(b2.weight);
}

bool fillFirst(BucketMap &map, const RawSequences &A, const RawSequences &B, int indexOffset,
               int &curBucket) {
  for (size_t i = 0; i < A.size(); ++i) {
    const size_t aLen = A[i].size(), bLen = B[i].size();
    const int w = search_from(map, curBucket, aLen + bLen);
    if (w < 0) return false;
    curBucket = w;
    place_ab(map.buckets[w], indexOffset + (int)i, aLen, bLen);
  }
  return true;
}

bool roundRobin(BucketMap &map, const RawSequences &A, const RawSequences &B, int indexOffset,
                int &curBucket) {
  for (size_t i = 0; i < A.size(); ++i) {
    const size_t aLen = A[i].size(), bLen = B[i].size();
    const int w = search_from(map, curBucket, aLen + bLen);
    if (w < 0) return false;
    BucketMapping &b = map.buckets[w];
    place_ab(b, indexOffset + (int)i, aLen, bLen);
    add_weight(b, aLen * bLen);
    curBucket = w + 1;
  }
  return true;
}

bool greedy(BucketMap &map, const RawSequences &A, const RawSequences &B, int indexOffset,
            BucketHeap &q) {
  const auto srts = by_cells(A.size(), [&](size_t i) { return A[i].size() * B[i].size(); });
  for (const auto &e : srts) {
    const int k = e.second;
    const size_t aLen = A[k].size(), bLen = B[k].size();
    BucketMapping *b = take_lightest(q, [&](const BucketMapping &x) {
      return fits(x, map, aLen + bLen);
    });
    if (!b) return false;
    place_ab(*b, indexOffset + k, aLen, bLen);
    add_weight(*b, aLen * bLen);
    q.push(std::ref(*b));
  }
  return true;
}

void fillFirst(BucketMap &map, const RawSequences &A, const RawSequences &B, int indexOffset) {
  int cur = 0;
  if (!fillFirst(map, A, B, indexOffset, cur)) throw std::runtime_error("Out of buckets.");
}

void roundRobin(BucketMap &map, const RawSequences &A, const RawSequences &B, int indexOffset) {
  int cur = 0;
  if (!roundRobin(map, A, B, indexOffset, cur)) throw std::runtime_error("Out of buckets.");
}

void greedy(BucketMap &map, const RawSequences &A, const RawSequences &B, int indexOffset) {
  BucketHeap q;
  for (auto &b : map.buckets) q.push(std::ref(b));
  if (!greedy(map, A, B, indexOffset, q)) throw std::runtime_error("Out of buckets.");
}

void fillFirst(BucketMap &map, const RawSequences &Seqs, const Comparisons &Cmps,
               int indexOffset) {
  int cur = 0;
  for (size_t i = 0; i < Cmps.size(); ++i) {
    const Comparison &c = Cmps[i];
    Need nd{(int)Seqs[c.indexA].size(), (int)Seqs[c.indexB].size()};
    const int w = next_bucket_mn(map, cur, c, nd, indexOffset);
    cur = w;
    place_mn(map.buckets[w], nd, c, indexOffset + (int)i, indexOffset);
  }
}

void roundRobin(BucketMap &map, const RawSequences &Seqs, const Comparisons &Cmps,
                int indexOffset) {
  int cur = 0;
  for (size_t i = 0; i < Cmps.size(); ++i) {
    const Comparison &c = Cmps[i];
    Need nd{(int)Seqs[c.indexA].size(), (int)Seqs[c.indexB].size()};
    const int w = next_bucket_mn(map, cur, c, nd, indexOffset);
    place_mn(map.buckets[w], nd, c, indexOffset + (int)i, indexOffset);
    cur = w + 1;
  }
}

void greedy(BucketMap &map, const RawSequences &Seqs, const Comparisons &Cmps, int indexOffset) {
  BucketHeap q;
  for (auto &b : map.buckets) q.push(std::ref(b));
  const auto srts = by_cells(Cmps.size(), [&](size_t i) {
    return Seqs[Cmps[i].indexA].size() * Seqs[Cmps[i].indexB].size();
  });
  for (const auto &e : srts) {
    const int k = e.second;
    const Comparison &c = Cmps[k];
    const size_t aLen = Seqs[c.indexA].size(), bLen = Seqs[c.indexB].size();
    Need nd{(int)aLen, (int)bLen};
    // the presence check compares the stored index with the plain sequence
    // index, as src/partition.cpp:318-327 does
    BucketMapping *b = take_lightest(q, [&](const BucketMapping &x) {
      Need t = nd;
      lookup(x, c.indexA, c.indexB, t);
      if (x.cmps.size() + 1 > (size_t)map.cmpCapacity ||
          (int64_t)x.seqSize + t.effA + t.effB > (int64_t)map.sequenceCapacity)
        return false;
      nd = t;
      return true;
    });
    if (!b) throw std::runtime_error("Out of buckets");
    place_mn(*b, nd, c, indexOffset + k, indexOffset);
    add_weight(*b, aLen * bLen);
    q.push(std::ref(*b));
  }
}

bool fillBuckets(Algorithm algo, BucketMap &map, const RawSequences &A, const RawSequences &B,
                 int indexOffset, int &curBucket, BucketHeap &heap) {
  switch (algo) {
    case Algorithm::fillFirst: return fillFirst(map, A, B, indexOffset, curBucket);
    case Algorithm::roundRobin: return roundRobin(map, A, B, indexOffset, curBucket);
    case Algorithm::greedy: return greedy(map, A, B, indexOffset, heap);
  }
  return false;
}

}  // namespace partition
}  // namespace ipu
