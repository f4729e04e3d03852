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
// Work array header: ISX_CLASSES claim counters, one per 128-byte line
// (words 32 c), then the segment total, the flags and the call epoch; the
// pairs' first segments start at word ISX_HDR.
constexpr uint32_t ISX_CLASSES = 8;
constexpr uint32_t ISX_CLAIM = 0;
constexpr uint32_t ISX_TOTAL = 32 * ISX_CLASSES;
constexpr uint32_t ISX_FLAGS = ISX_TOTAL + 1;
constexpr uint32_t ISX_EPOCH = ISX_TOTAL + 2;
constexpr uint32_t ISX_HDR = ISX_TOTAL + 64;

// Segments of a pair with list lengths la, lb (none when a list is empty).
__host__ __device__ constexpr uint64_t isx_segments(uint32_t la, uint32_t lb) {
    const uint64_t na = la < lb ? la : lb, nb = la < lb ? lb : la;
    if (na == 0) return 0;
    if (nb >= (uint64_t)ISX_SKEW * na) return (na + ISX_CHUNK - 1) / ISX_CHUNK;
    return (na + nb + ISX_SEG - 1) / ISX_SEG;
}

// First word of the per-segment status array (64-bit aligned) after the
// header and the pairs' first-segment table.
__host__ __device__ constexpr uint64_t isx_status_word(uint32_t npairs) {
    return (ISX_HDR + (uint64_t)npairs + 1 + 1) & ~1ull;
}

// Words of the work array for npairs pairs and cap segments.
__host__ __device__ constexpr uint64_t isx_work_words(uint32_t npairs, uint64_t cap) {
    return isx_status_word(npairs) + 2 * cap;
}

// Work array flags (word ISX_FLAGS).
constexpr uint32_t ISX_OVER = 1u;      // the segments exceed the array: no work done
constexpr uint32_t ISX_GAVE_UP = 2u;   // a look-back wait passed its limit

// Launches the plan and main kernels on `stream`; allocates nothing. `work`
// holds isx_work_words(npairs, cap) words, zeroed before its first use (the
// status words carry a call epoch kept in the header, so later calls need no
// clearing). The main kernel runs the blocks that fit at once on `cus`
// compute units.
void isx_launch(hipStream_t stream, uint32_t cus, uint32_t *work, uint32_t cap, const uint32_t *data,
                const uint64_t *a_off, const uint32_t *a_len, const uint64_t *b_off, const uint32_t *b_len,
                uint32_t npairs, uint32_t *counts, const uint64_t *out_off, uint32_t *out);

// Resident blocks per CU of the main kernel, with output or count only.
int isx_blocks_per_cu(bool out);

}  // namespace fm
