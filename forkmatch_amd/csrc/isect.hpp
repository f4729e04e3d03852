// Batched intersection of sorted u32 lists on the device (isect.hip).
// Shared by the engine's C ABI (fm_intersect_pairs and its no-allocation
// variant) and nothing else.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fm {

// Merge-path elements per merge segment.
constexpr uint32_t ISX_SEG = 16384;
// Elements of the shorter list per skew segment (a pair whose longer list is
// at least ISX_SKEW times the shorter one).
constexpr uint32_t ISX_CHUNK = 512;
constexpr uint32_t ISX_SKEW = 32;
// Work array: a header of ISX_HDR words (ISX_CLASSES claim counters, one per
// 128-byte line at words 32 c, then the segment total, the flags and the
// call epoch), one 64-bit status word per segment the array can hold (cap),
// then the pairs' first segments (npairs + 1 words). A work array keeps its
// cap while it lives, so every status word of an earlier call on it is a
// status word of an earlier epoch, never another table's data.
constexpr uint32_t ISX_CLASSES = 8;
constexpr uint32_t ISX_CLAIM = 0;
constexpr uint32_t ISX_TOTAL = 32 * ISX_CLASSES;
constexpr uint32_t ISX_FLAGS = ISX_TOTAL + 1;
constexpr uint32_t ISX_EPOCH = ISX_TOTAL + 2;
constexpr uint32_t ISX_HDR = ISX_TOTAL + 64;
static_assert(ISX_HDR % 2 == 0, "status words are 64-bit aligned");

// Segments of a pair with list lengths la, lb (none when a list is empty).
__host__ __device__ constexpr uint64_t isx_segments(uint32_t la, uint32_t lb) {
    const uint64_t na = la < lb ? la : lb, nb = la < lb ? lb : la;
    if (na == 0) return 0;
    if (nb >= (uint64_t)ISX_SKEW * na) return (na + ISX_CHUNK - 1) / ISX_CHUNK;
    return (na + nb + ISX_SEG - 1) / ISX_SEG;
}

// First word of the pairs' first-segment table in an array of cap segments.
__host__ __device__ constexpr uint64_t isx_table_word(uint64_t cap) { return ISX_HDR + 2 * cap; }

// Words of a work array for up to pairs - 1 pairs and cap segments.
__host__ __device__ constexpr uint64_t isx_work_words(uint64_t pairs, uint64_t cap) {
    return isx_table_word(cap) + pairs;
}

// Work array flags (word ISX_FLAGS).
constexpr uint32_t ISX_OVER = 1u;      // the segments exceed the array: no work done
constexpr uint32_t ISX_GAVE_UP = 2u;   // a look-back wait passed its limit

// Launches the plan and main kernels on `stream`; allocates nothing. `work`
// holds at least isx_work_words(npairs + 1, cap) words, zeroed before its
// first use, and is always used with the same cap (the status words carry a
// call epoch kept in the header, so later calls need no clearing). The main
// kernel runs the blocks that fit at once on `cus` compute units.
void isx_launch(hipStream_t stream, uint32_t cus, uint32_t *work, uint32_t cap, const uint32_t *data,
                const uint64_t *a_off, const uint32_t *a_len, const uint64_t *b_off, const uint32_t *b_len,
                uint32_t npairs, uint32_t *counts, const uint64_t *out_off, uint32_t *out);

// Resident blocks per CU of the main kernel, with output or count only.
int isx_blocks_per_cu(bool out);

}  // namespace fm
