// Size limits of the device data structures.  No HIP dependencies, so the
// host-side limit test compiles this header with g++ alone.
#pragma once
#include <stdint.h>

namespace srt {

// The 4-wide traversal reads its nodes through a buffer resource: the
// resource's size field and the per-lane byte offsets are 32-bit, signed in
// the size computation, so every byte of the node array must lie below 2^31.
// A tree with more 4-wide nodes is traversed in its binary form (64-bit
// addressing), which gives the same hits.
constexpr int64_t kNode4Bytes = 128;
constexpr int64_t kMaxNodes4 = 0x7fffffffll / kNode4Bytes;   // 16,777,215 nodes
inline bool nodes4_fit(int64_t nodes4, int64_t limit = kMaxNodes4) {
    return nodes4 >= 0 && nodes4 <= (limit < kMaxNodes4 ? limit : kMaxNodes4);
}

// Leaf references pack the first sorted triangle into 28 bits (bit 31 marks a
// leaf, bits 0-2 hold the count).
constexpr int64_t kMaxTriangles = (1ll << 28) - 1;

// Path slots of a chunk are 32-bit indices (queue entries, slot offsets of the
// segments); fewer than one wave of slots is not useful.
constexpr int64_t kMinPathPool = 64;
constexpr int64_t kMaxPathPool = 0x7fffffffll;
inline bool path_pool_ok(int64_t slots) { return slots >= kMinPathPool && slots <= kMaxPathPool; }

}  // namespace srt
