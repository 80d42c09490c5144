// The device view of a join table and its lookup, shared by join.hip,
// aggregate.hip and the aggregation kernels generated at run time; compiles
// under hipcc and hipRTC.
//
// The table indexes the build side's rows grouped by key, build order within
// a key: each distinct non-NULL key has a run [start, start+count) in `rows`
// (the layout of the reference's JoinTable, crates/oxidelake-device/src/cpu/
// kernels.rs:181-266). Two forms, picked at build time from the key range:
//  - dense: keys in [min, min+range) address their slot directly (no hashing,
//    one read per probe) when the range is at most twice the build rows;
//  - hashed: open addressing, capacity a power of two at least twice the
//    build rows, 64-bit keys claimed by one CAS (INT64_MIN marks an empty
//    slot; the key INT64_MIN itself has the extra slot `cap`).
#pragma once

#include "scalar.h"

namespace oxr {

constexpr int64_t kJoinEmpty = kI64Min;
constexpr uint64_t kNoSlot = ~(uint64_t)0;

struct JTable {
  int32_t dense;
  int32_t unique;     // every key has exactly one build row
  int64_t min_key;    // dense: smallest key
  uint64_t slots;     // dense: range; hashed: cap (power of two); +1 slot follows
  int64_t *keys;      // hashed only: slots+1
  uint32_t *start;    // slots+1
  uint32_t *count;    // slots+1
  uint32_t *rows;     // build rows with a non-NULL key, grouped by slot
};

__host__ __device__ inline uint64_t join_hash(uint64_t h) {
  // MurmurHash3 finalizer
  h ^= h >> 33;
  h *= 0xff51afd7ed558ccdull;
  h ^= h >> 33;
  h *= 0xc4ceb9fe1a85ec53ull;
  h ^= h >> 33;
  return h;
}

// Slot of `key`, or kNoSlot when the table does not hold it.
__device__ inline uint64_t join_find(const JTable &t, int64_t key) {
  if (t.dense) {
    const uint64_t d = (uint64_t)key - (uint64_t)t.min_key;
    if (key < t.min_key || d >= t.slots) return kNoSlot;
    return t.count[d] ? d : kNoSlot;
  }
  if (key == kJoinEmpty) return t.count[t.slots] ? t.slots : kNoSlot;
  const uint64_t mask = t.slots - 1;
  uint64_t h = join_hash((uint64_t)key) & mask;
  for (uint64_t p = 0; p < t.slots; p++) {
    const int64_t k = t.keys[h];
    if (k == key) return h;
    if (k == kJoinEmpty) return kNoSlot;
    h = (h + 1) & mask;
  }
  return kNoSlot;
}

}  // namespace oxr
