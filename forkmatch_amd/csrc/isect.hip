// Wave64 intersection of pairs of strictly increasing u32 lists: counts and,
// optionally, every intersection written in ascending order. Replaces the
// warp-cooperative intersect<CoopSize,Decomp> and its binary searches
// (src/intersect.cuh:31-191); the reference's intended use is the candidate
// intersection of isoGen (src/smsm.cu:297-345).
//
// A pair (A = the shorter list) is cut into segments:
//  - merge segments: ISX_SEG consecutive elements of the merge path of A and
//    B (the two lists merged, A first on ties). A segment holds A[i0, i1)
//    and B[j0, j1) plus B[j1], the one later B element that can equal A[i1-1];
//    every match is found in the segment that holds its A element;
//  - skew segments (the longer list at least ISX_SKEW times the shorter):
//    ISX_CHUNK elements of A, looked up in B.
// One wave works on one segment at a time, claimed from sharded counters
// (fm_isx_kernel explains the order and why every wait ends):
//  1. its pair by a 64-ary search of the pairs' first segments, and its cuts
//     (the two merge-path splits, one per half wave, 32-ary: four rounds of
//     loads for lists of 10^5 elements);
//  2. dense segments (A part values spanning fewer than BMSPAN = 32,768
//     integers): the A part and then the B part are streamed in 512-element
//     chunks, four in flight per wave (16-byte buffer loads), into two LDS
//     bitmaps over the A part's range (A's elements; B's elements in that
//     range); the count is the popcount of their AND and the matches are the
//     values of its set bits. Sparser segments hold one A chunk at a time
//     sorted in LDS and search each B element there (9 LDS reads), advancing
//     A chunks and B windows by their last values as a merge does; skew
//     segments find the B range of their A chunk (two 32-ary searches) and
//     binary-search each A element in it;
//  3. the number of matches in the pair's earlier segments comes from
//     decoupled look-back over per-segment status words (the segment's count
//     is published first, then the inclusive sum), and the matches are
//     stored at that offset with coalesced stores; a dense segment's are
//     written after the wave has built its next segment, so the look-back
//     mostly finds the earlier counts ready. Count-only calls add each
//     segment's count to the pair's count instead.
// Every element of a dense merge segment is read from HBM once; a sparse
// segment with output reads its lists twice (to count, then to write).
#include "isect.hpp"

#include <algorithm>

namespace fm {
namespace {

constexpr uint32_t W = 64;
constexpr uint32_t K = 8;                // registers per lane of a chunk
constexpr uint32_t CH = K * W;           // elements per chunk
constexpr uint32_t P = 4;                // chunks in flight per wave (dense segments)
constexpr uint32_t BMW = 1024;           // LDS words of a dense segment's bitmap
constexpr uint32_t BMSPAN = BMW * 32;    // dense: the A part spans fewer values
constexpr uint32_t WPB_COUNT = 2;        // waves per block, count only
constexpr uint32_t NB = 2;               // builder waves per block with output (and one emitter)
constexpr uint32_t NS = 4;               // handoff slots per block with output
constexpr uint32_t NE = 1;               // emitter waves per block with output
// Count-only blocks per CU: more concurrent streams read slower (measured:
// 8 waves per CU beat 10, 12 and 16)
constexpr int COUNT_BLOCKS_PER_CU = 4;
constexpr uint32_t RUN = 16;             // consecutive segments per claim-class run
// A look-back wait gives up after this long (s_memrealtime runs at 100 MHz).
constexpr unsigned long long WAIT_TICKS = 200000000ull;
static_assert(ISX_CHUNK == CH, "a skew segment is one chunk");

// Status word of a segment: bits 0-31 a count, bit 32 the segment's own
// count, bit 33 the sum over the pair's segments up to and including it,
// bits 34-63 the call's epoch (words of earlier calls never read as ready).
constexpr unsigned long long ST_AGG = 1ull << 32, ST_INCL = 2ull << 32;

__device__ __forceinline__ uint32_t lane() { return __lane_id(); }
__device__ __forceinline__ uint32_t rdl(uint32_t v, uint32_t l) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)l);
}
__device__ __forceinline__ uint32_t ld_rlx(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t mbcnt(unsigned long long m, uint32_t base) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, base));
}
// v moved across lanes by a DPP control (0 where a lane has no source or
// its row is masked off).
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ uint32_t dpp(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xf, false);
}
// Inclusive prefix sum over the wave: shifts by 1, 2, 4, 8 lanes within each
// row of 16, then row 0's total into row 1 and row 2's into row 3
// (row_bcast:15), then rows 0-1's total into rows 2-3 (row_bcast:31).
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    v += dpp<0x111>(v);
    v += dpp<0x112>(v);
    v += dpp<0x114>(v);
    v += dpp<0x118>(v);
    v += dpp<0x142, 0xa>(v);
    v += dpp<0x143, 0xc>(v);
    return v;
}
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) { return rdl(wave_incl_scan(v), W - 1); }

// A load of a value that no kernel of the call writes while it reads it,
// through the constant address space: with a uniform address it becomes a
// scalar load (tracked apart from the vector loads in flight).
template <class T>
__device__ __forceinline__ T ldc(const T *p) {
    return *(const __attribute__((address_space(4))) T *)p;
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
// four words at a 4-byte-aligned address (global accesses of several dwords
// need only dword alignment)
typedef uint32_t u32x4_a4 __attribute__((ext_vector_type(4), aligned(4)));

struct Pair {
    const uint32_t *A, *B;
    uint32_t na, nb;
};

__device__ __forceinline__ Pair pair_of(const uint32_t *data, const uint64_t *a_off, const uint32_t *a_len,
                                        const uint64_t *b_off, const uint32_t *b_len, uint32_t p) {
    p = __builtin_amdgcn_readfirstlane(p);
    Pair r{data + ldc(a_off + p), data + ldc(b_off + p), ldc(a_len + p), ldc(b_len + p)};
    if (r.na > r.nb) {
        const uint32_t *t = r.A;
        r.A = r.B;
        r.B = t;
        const uint32_t n = r.na;
        r.na = r.nb;
        r.nb = n;
    }
    return r;
}

// Largest p < npairs with seg_begin[p] <= g (the pair of segment g): 64
// samples per round.
__device__ __forceinline__ uint32_t find_pair(const uint32_t *seg_begin, uint32_t npairs, uint32_t g) {
    uint32_t lo = 0, hi = npairs;  // seg_begin[lo] <= g; hi == npairs or seg_begin[hi] > g
    while (hi - lo > 1) {
        const uint32_t step = (hi - lo + W - 1) / W;
        const uint32_t idx = lo + lane() * step;
        const bool pr = lane() * step < hi - lo && seg_begin[idx] <= g;
        const uint32_t t = (uint32_t)__popcll(__ballot(pr));  // >= 1: idx == lo on lane 0
        hi = min(lo + t * step, hi);
        lo = lo + (t - 1) * step;
    }
    return lo;
}

// Per half wave (lanes 0-31 and 32-63, each with its own lo, hi and
// predicate): the first index in [lo, hi) where pred is false, or hi;
// pred must be true then false over the range. 32 samples per round.
template <class Pred>
__device__ __forceinline__ uint32_t first_false2(uint64_t lo, uint64_t hi, Pred pred) {
    const uint32_t k = lane() & 31;
    const bool upper = lane() >= 32;
    while (__ballot(lo < hi)) {
        const uint64_t len = hi - lo, step = (len + 31) >> 5;
        const bool probe = lo < hi && k * step < len;
        const bool pr = probe && pred(lo + k * step);
        const unsigned long long m = __ballot(pr);
        const uint32_t t = (uint32_t)__popc(upper ? (uint32_t)(m >> 32) : (uint32_t)m);
        if (lo < hi) {
            if (t == 0) {
                hi = lo;
            } else {
                const uint64_t nlo = lo + (t - 1) * step + 1;
                hi = min(lo + t * step, hi);
                lo = nlo;
            }
        }
    }
    return (uint32_t)lo;
}

// One chunk of n >= 1 valid elements from p: register k of lane l holds
// element k * 64 + l, clamped to the last valid one.
__device__ __forceinline__ void load_chunk(const uint32_t *p, uint32_t n, uint32_t (&r)[K]) {
    const uint32_t l = lane();
    if (n >= CH) {
#pragma unroll
        for (uint32_t k = 0; k < K; k++) r[k] = p[k * W + l];
    } else {
        const uint32_t last = n - 1;
#pragma unroll
        for (uint32_t k = 0; k < K; k++) r[k] = p[min(k * W + l, last)];
    }
}

// Adds x to the wave's matches where m holds; with WRITE the matches go to
// o[c + rank] (rank = matching lanes below, so in lane order), below lim.
template <bool WRITE>
__device__ __forceinline__ uint32_t emit(bool m, uint32_t x, uint32_t c, uint32_t *o, uint32_t lim) {
    const unsigned long long mk = __ballot(m);
    if (WRITE) {
        const uint32_t pos = mbcnt(mk, c);
        if (m && pos < lim) o[pos] = x;
    }
    return c + (uint32_t)__popcll(mk);
}

// Cache policy of the streamed chunk loads: non-temporal (each list element
// is read once; keeping them out of the way of the L2's other lines measured
// 8% faster).
constexpr int AUX_NT = 2;

// One chunk of n elements from p for the dense path, two 16-byte loads per
// lane through a buffer resource: lane l holds elements 4l..4l+3 in r[0..3]
// and 256+4l..256+4l+3 in r[4..7]; elements at or past n read as 0 (the
// buffer's range check, per dword; nothing is fetched for them).
__device__ __forceinline__ void load_chunk_x4(const uint32_t *p, uint32_t n, uint32_t (&r)[K]) {
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t *>(p), (short)0, (int)(min(n, CH) * 4), 0x00020000);
    const int voff = (int)(lane() * 16);
    const u32x4 a = __builtin_amdgcn_raw_buffer_load_b128(rs, voff, 0, AUX_NT);
    const u32x4 b = __builtin_amdgcn_raw_buffer_load_b128(rs, voff + (int)(CH / 2 * 4), 0, AUX_NT);
    r[0] = a.x, r[1] = a.y, r[2] = a.z, r[3] = a.w;
    r[4] = b.x, r[5] = b.y, r[6] = b.z, r[7] = b.w;
}

// Sets the bits of one chunk's elements (load_chunk_x4 layout) in the bitmap
// of [amin, amin + span]. A lane's four consecutive elements of a group fall
// in the word of the group's first element and the next one unless the list
// is sparse there, so each group takes two atomic ORs (neighbouring lanes
// share at most a word) instead of one per element; other elements are set
// one by one. A chunks lie in the range; B chunks set only their elements in
// it (CHECK). Zeros read past a partial chunk's end set nothing (PARTIAL).
template <bool PARTIAL, bool CHECK>
__device__ __forceinline__ void set_bits(uint32_t *bm, const uint32_t (&r)[K], uint32_t amin, uint32_t span,
                                         uint32_t n) {
    const uint32_t l = lane();
#pragma unroll
    for (uint32_t h = 0; h < 2; h++) {
        uint32_t off[4];
        bool ok[4];
#pragma unroll
        for (uint32_t e = 0; e < 4; e++) {
            off[e] = r[4 * h + e] - amin;
            ok[e] = true;
            if (CHECK) ok[e] = off[e] <= span;
            if (PARTIAL) ok[e] = ok[e] & (h * (CH / 2) + 4 * l + e < n);
        }
        const uint32_t wb = off[0] >> 5;  // the group's first word (valid when ok[0])
        uint32_t lo = 0, hi = 0;
        bool rare[4];
#pragma unroll
        for (uint32_t e = 0; e < 4; e++) {
            const uint32_t rel = off[e] - (wb << 5);
            const bool near = ok[0] && rel < 64;
            const uint32_t bit = 1u << (rel & 31);
            lo |= ok[e] && near && rel < 32 ? bit : 0u;
            hi |= ok[e] && near && rel >= 32 ? bit : 0u;
            rare[e] = ok[e] && !near;
        }
        if (ok[0]) {
            atomicOr(&bm[wb & (BMW - 1)], lo);
            if (hi) atomicOr(&bm[(wb + 1) & (BMW - 1)], hi);
        }
        if (__ballot(rare[0] | rare[1] | rare[2] | rare[3])) {
#pragma unroll
            for (uint32_t e = 0; e < 4; e++)
                if (rare[e]) atomicOr(&bm[(off[e] >> 5) & (BMW - 1)], 1u << (off[e] & 31));
        }
    }
}

// Dense merge segment: A part pa[0, la) with values in [amin, amin + span],
// span < BMSPAN, and B part pb[0, lb). Builds the bitmap of the A part (bmA)
// and of the B elements in its range (bmB), both span / 32 + 1 words. The A
// chunks, then the B chunks, are streamed P at a time: chunk c + P is loaded
// into the registers chunk c leaves (chunks past the segment load nothing),
// so every load is issued P chunks before its use and the load sequence is
// the same on every path.
__device__ __forceinline__ void dense_build(const uint32_t *pa, uint32_t la, const uint32_t *pb, uint32_t lb,
                                            uint32_t amin, uint32_t span, uint32_t *bmA, uint32_t *bmB) {
    const uint32_t l = lane();
    const uint32_t nA = (la + CH - 1) / CH, nC = nA + (lb + CH - 1) / CH;
    const uint32_t nw = la ? (span >> 5) + 1 : 0u;
    for (uint32_t w = l; w < nw; w += W) {
        bmA[w] = 0;
        bmB[w] = 0;
    }
    // chunk c: its first element and its number of elements (0 past the segment)
    auto chunk = [&](uint32_t c, uint32_t(&r)[K]) {
        const bool in_a = c < nA;
        const uint32_t *const p = in_a ? pa + c * CH : pb + (c - nA) * CH;
        const uint32_t n = in_a ? la - c * CH : (c < nC ? lb - (c - nA) * CH : 0u);
        load_chunk_x4(p, n, r);
    };
    uint32_t r[P][K];
#pragma unroll
    for (uint32_t s = 0; s < P; s++) chunk(s, r[s]);
    for (uint32_t c0 = 0; c0 < nC; c0 += P) {
#pragma unroll
        for (uint32_t s = 0; s < P; s++) {
            const uint32_t cc = c0 + s;
            if (cc < nA) {
                const uint32_t n = la - cc * CH;
                if (n >= CH) set_bits<false, false>(bmA, r[s], amin, span, n);
                else set_bits<true, false>(bmA, r[s], amin, span, n);
            } else if (cc < nC && la) {
                const uint32_t n = lb - (cc - nA) * CH;
                if (n >= CH) set_bits<false, true>(bmB, r[s], amin, span, n);
                else set_bits<true, true>(bmB, r[s], amin, span, n);
            }
            chunk(cc + P, r[s]);
        }
    }
}

// The number of common bits of the nw words of bmA and bmB; with WRITE the
// common bits are also written to dst (which may be bmA).
template <bool WRITE>
__device__ __forceinline__ uint32_t common_bits(uint32_t *bmA, const uint32_t *bmB, uint32_t nw, uint32_t *dst) {
    uint32_t c = 0;
    for (uint32_t w = lane(); w < nw; w += W) {
        const uint32_t x = bmA[w] & bmB[w];
        if (WRITE) dst[w] = x;
        c += (uint32_t)__popc(x);
    }
    return wave_sum(c);
}

// Writes the values of the bits of bm (nw words over [amin, ...]) in
// ascending order to o[0, lim): 64 words per round; each
// lane puts its word's values into the LDS stage (CH words) at the word's
// offset (a wave prefix sum of the bit counts), and the stage leaves with
// coalesced stores.
__device__ __forceinline__ void dense_emit(const uint32_t *bm, uint32_t nw, uint32_t amin, uint32_t *stage,
                                           uint32_t *o, uint32_t lim) {
    const uint32_t l = lane();
    uint32_t base = 0;
    for (uint32_t w0 = 0; w0 < nw; w0 += W) {
        const uint32_t w = w0 + l;
        const uint32_t bits = w < nw ? bm[w] : 0u;
        const uint32_t n = (uint32_t)__popc(bits);
        const uint32_t incl = wave_incl_scan(n);
        const uint32_t tot = rdl(incl, W - 1);
        const uint32_t v0 = amin + 32 * w;
        for (uint32_t lo = 0; lo < tot; lo += CH) {
            uint32_t b = bits, pos = incl - n;
            while (b) {
                if (pos >= lo && pos < lo + CH) stage[pos - lo] = v0 + (uint32_t)__builtin_ctz(b);
                pos++;
                b &= b - 1;
            }
            // the stage leaves in non-temporal 16-byte stores (whole groups
            // of four), the group at its end or at lim dword by dword
            const uint32_t cnt = min(CH, tot - lo), at0 = base + lo;
            for (uint32_t i = 4 * l; i < cnt; i += 4 * W) {
                const u32x4 v = *reinterpret_cast<const u32x4 *>(stage + i);
                const uint32_t at = at0 + i;
                if (i + 4 <= cnt && at + 4 <= lim) {
                    __builtin_nontemporal_store(v, reinterpret_cast<u32x4_a4 *>(o + at));
                } else {
#pragma unroll
                    for (uint32_t e = 0; e < 4; e++)
                        if (i + e < cnt && at + e < lim) o[at + e] = v[e];
                }
            }
        }
        base += tot;
    }
}

// Sparse merge segment: one A chunk at a time sorted in LDS (cs, CH words),
// one B window in registers; each B element in the chunk's range is found
// by a 9-step search of the chunk. The chunk or the window that ends first
// (by last value) is replaced, both on a tie, so every chunk meets every
// window whose values overlap its own. Returns the number of matches; with
// WRITE they are also stored at o[0, lim).
template <bool WRITE>
__device__ __forceinline__ uint32_t sparse_segment(const uint32_t *pa, uint32_t la, const uint32_t *pb, uint32_t lb,
                                                   uint32_t *cs, uint32_t *o, uint32_t lim) {
    const uint32_t l = lane();
    uint32_t a[K], b[K];
    uint32_t ia = 0, jb = 0, c = 0, amin = 0, amax = 0, bmax = 0;
    auto next_a = [&]() {
        load_chunk(pa + ia, la - ia, a);
        amin = rdl(a[0], 0);
        amax = rdl(a[K - 1], W - 1);
#pragma unroll
        for (uint32_t k = 0; k < K; k++) cs[k * W + l] = a[k];
    };
    auto next_b = [&]() {
        load_chunk(pb + jb, lb - jb, b);
        bmax = rdl(b[K - 1], W - 1);
    };
    bool a_has = la > 0, b_has = lb > 0;
    if (a_has) next_a();
    if (b_has) next_b();
    while (a_has || b_has) {
        if (a_has && b_has) {
            const uint32_t n = lb - jb;
#pragma unroll
            for (uint32_t k = 0; k < K; k++) {
                const uint32_t x = b[k];
                const bool in = x >= amin && x <= amax && (n >= CH || k * W + l < n);
                uint32_t pos = 0;  // lower bound of x among the chunk's 512 words (the last one is amax)
#pragma unroll
                for (uint32_t h = CH / 2; h >= 1; h >>= 1) pos = cs[pos + h - 1] < x ? pos + h : pos;
                c = emit<WRITE>(in && cs[pos] == x, x, c, o, lim);
            }
        }
        bool adv_a = a_has, adv_b = b_has;
        if (a_has && b_has) {
            adv_b = bmax <= amax;
            adv_a = amax <= bmax;
        }
        if (adv_b) {
            jb += CH;
            b_has = jb < lb;
            if (b_has) next_b();
        }
        if (adv_a) {
            ia += CH;
            a_has = ia < la;
            if (a_has) next_a();
        }
    }
    return c;
}

// Skew segment: A chunk pa[0, la), la <= CH, looked up in all of B: its
// elements in a, found[k] where a[k] is in B. Returns the number found.
__device__ __forceinline__ uint32_t skew_find(const uint32_t *pa, uint32_t la, const uint32_t *B, uint32_t nb,
                                              uint32_t (&a)[K], bool (&found)[K]) {
    const uint32_t l = lane();
    load_chunk(pa, la, a);
    const uint32_t amin = rdl(a[0], 0), amax = rdl(a[K - 1], W - 1);
    // B's elements in [amin, amax]: lower bound of amin (lanes 0-31), upper bound of amax (32-63)
    const bool upper = l >= 32;
    const uint32_t x = upper ? amax : amin;
    const uint32_t r = first_false2(0, nb, [&](uint64_t j) { return upper ? B[j] <= x : B[j] < x; });
    const uint32_t lo = rdl(r, 0), hi = rdl(r, 32);
    uint32_t c = 0;
#pragma unroll
    for (uint32_t k = 0; k < K; k++) found[k] = false;
    if (lo < hi) {
        // base[k]: the last position in [lo, hi) whose element is <= a[k], or lo
        uint32_t base[K];
#pragma unroll
        for (uint32_t k = 0; k < K; k++) base[k] = lo;
        for (uint32_t n = hi - lo; n > 1;) {
            const uint32_t h = n >> 1;
#pragma unroll
            for (uint32_t k = 0; k < K; k++) base[k] = B[base[k] + h] <= a[k] ? base[k] + h : base[k];
            n -= h;
        }
#pragma unroll
        for (uint32_t k = 0; k < K; k++) {
            found[k] = k * W + l < la && B[base[k]] == a[k];
            c += (uint32_t)__popcll(__ballot(found[k]));
        }
    }
    return c;
}

// Publishes segment g's count c: the inclusive sum already when g is its
// pair's first segment (sb).
__device__ __forceinline__ void publish(unsigned long long *status, uint32_t g, uint32_t sb, uint32_t epoch,
                                        uint32_t c) {
    const unsigned long long tag = (unsigned long long)epoch << 34;
    if (lane() == 0)
        __hip_atomic_store(&status[g], tag | (g == sb ? ST_INCL : ST_AGG) | c, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// The count of the pair's segments before g (its count c published), by
// decoupled look-back over the earlier segments' status words; then g's
// inclusive sum is published. Waits only on segments claimed before g.
__device__ __forceinline__ uint32_t resolve(unsigned long long *status, uint32_t g, uint32_t sb, uint32_t epoch,
                                            uint32_t c, uint32_t *work) {
    if (g == sb) return 0;
    const unsigned long long tag = (unsigned long long)epoch << 34;
    const uint32_t l = lane();
    uint32_t prefix = 0, t = g - 1;  // lane k reads segment t - k
    const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
    while (true) {
        const bool in = l <= t - sb;
        const unsigned long long st =
            in ? __hip_atomic_load(&status[t - l], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : (tag | ST_INCL);
        const bool ready = (uint32_t)(st >> 34) == epoch && (st & (ST_AGG | ST_INCL)) != 0;
        const unsigned long long incl = __ballot(ready && (st & ST_INCL)), waiting = __ballot(!ready);
        if (incl) {
            const uint32_t f = (uint32_t)__ffsll((long long)incl) - 1;  // the nearest inclusive sum
            const unsigned long long upto = f == 63 ? ~0ull : ((2ull << f) - 1);
            if ((waiting & upto) == 0) {
                prefix += wave_sum(l <= f ? (uint32_t)st : 0u);
                break;
            }
        } else if (waiting == 0) {
            prefix += wave_sum((uint32_t)st);  // 64 published counts, none inclusive
            t -= W;
            continue;
        }
        if (__builtin_amdgcn_s_memrealtime() - t0 > WAIT_TICKS) {
            if (l == 0) atomicOr(&work[ISX_FLAGS], ISX_GAVE_UP);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    if (l == 0)
        __hip_atomic_store(&status[g], tag | ST_INCL | (uint32_t)(prefix + c), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    return prefix;
}

// One look-back attempt for segment g (its count c published) that does not
// wait: true with the count of the pair's segments before g in prefix (and
// g's inclusive sum published) when every segment it needs has published,
// false otherwise.
__device__ __forceinline__ bool try_resolve(unsigned long long *status, uint32_t g, uint32_t sb, uint32_t epoch,
                                            uint32_t c, uint32_t &prefix) {
    const unsigned long long tag = (unsigned long long)epoch << 34;
    const uint32_t l = lane();
    prefix = 0;
    if (g != sb) {
        uint32_t t = g - 1;  // lane k reads segment t - k
        while (true) {
            const bool in = l <= t - sb;
            const unsigned long long st =
                in ? __hip_atomic_load(&status[t - l], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : (tag | ST_INCL);
            const bool ready = (uint32_t)(st >> 34) == epoch && (st & (ST_AGG | ST_INCL)) != 0;
            const unsigned long long incl = __ballot(ready && (st & ST_INCL)), waiting = __ballot(!ready);
            if (incl) {
                const uint32_t f = (uint32_t)__ffsll((long long)incl) - 1;  // the nearest inclusive sum
                const unsigned long long upto = f == 63 ? ~0ull : ((2ull << f) - 1);
                if (waiting & upto) return false;
                prefix += wave_sum(l <= f ? (uint32_t)st : 0u);
                break;
            }
            if (waiting) return false;
            prefix += wave_sum((uint32_t)st);  // 64 published counts, none inclusive
            t -= W;
        }
        if (l == 0)
            __hip_atomic_store(&status[g], tag | ST_INCL | (uint32_t)(prefix + c), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
    return true;
}

// One block of 1,024 threads: every pair's first segment (a scan of the
// segment counts; thread t takes a run of consecutive pairs, so its length
// loads are issued together), the counts cleared, the claim counters
// cleared, the segment total (0 and the ISX_OVER flag when it exceeds cap),
// a new epoch. When the epoch wraps, the status words are cleared first, so
// no word of a call 2^30 calls back reads as current.
__global__ __launch_bounds__(1024) void fm_isx_plan_kernel(const uint32_t *a_len, const uint32_t *b_len,
                                                           uint32_t npairs, uint32_t *counts, uint32_t *work,
                                                           uint32_t cap) {
    constexpr uint32_t T = 1024, EPOCH_MASK = (1u << 30) - 1;
    __shared__ unsigned long long s_part[T / W];
    const uint32_t tid = threadIdx.x, l = lane(), wid = tid / W;
    uint32_t *const seg_begin = work + isx_table_word(cap);
    if (tid < ISX_CLASSES) work[ISX_CLAIM + 32 * tid] = 0;
    const uint32_t run = (npairs + T - 1) / T;
    const uint32_t r0 = min(tid * run, npairs), r1 = min(r0 + run, npairs);
    unsigned long long mine = 0;
    for (uint32_t p = r0; p < r1; p++) {
        mine += isx_segments(a_len[p], b_len[p]);
        counts[p] = 0;
    }
    unsigned long long v = mine;
#pragma unroll
    for (int o = 1; o < (int)W; o <<= 1) {
        const unsigned long long t = __shfl_up(v, o, W);
        if (l >= (uint32_t)o) v += t;
    }
    if (l == W - 1) s_part[wid] = v;
    const uint32_t old_epoch = work[ISX_EPOCH];
    if (old_epoch == EPOCH_MASK) {
        // the epoch wraps: clear the status words (every segment the array can hold)
        unsigned long long *const status = reinterpret_cast<unsigned long long *>(work + ISX_HDR);
        for (uint32_t g = tid; g < cap; g += T) status[g] = 0;
    }
    __syncthreads();
    unsigned long long begin = v - mine, total = 0;
    for (uint32_t w = 0; w < T / W; w++) {
        if (w < wid) begin += s_part[w];
        total += s_part[w];
    }
    for (uint32_t p = r0; p < r1; p++) {
        seg_begin[p] = begin <= cap ? (uint32_t)begin : cap;
        begin += isx_segments(a_len[p], b_len[p]);
    }
    if (tid == 0) {
        const bool over = total > cap;
        seg_begin[npairs] = over ? cap : (uint32_t)total;
        work[ISX_TOTAL] = over ? 0u : (uint32_t)total;
        work[ISX_FLAGS] = over ? ISX_OVER : 0u;
        const uint32_t ep = (old_epoch + 1) & EPOCH_MASK;
        work[ISX_EPOCH] = ep ? ep : 1u;
    }
}

// A dense segment handed from a builder wave to its block's emitter wave:
// its common bits sit in a slot of BMW words.
struct Handoff {
    uint32_t g, sb, p, nw, amin, c, last, room;
    uint64_t off;
    unsigned long long since;  // when an emitter first found its look-back waiting (0: not yet)
};
constexpr uint32_t SLOT_FREE = 0, SLOT_FILLING = 1, SLOT_READY = 2, SLOT_EMITTING = 3;

// A block's LDS. Count only: WPB_COUNT builder areas (two bitmaps, then CH
// words for a sparse segment's sorted A chunk or a stage). With output: NB
// builder areas, NS slots with their handoffs and states, the number of
// builders that have finished, and a stage per emitter.
template <bool OUT>
struct BlockLds {
    uint32_t builder[WPB_COUNT][2 * BMW + CH];
};
template <>
struct BlockLds<true> {
    uint32_t builder[NB][2 * BMW + CH];
    uint32_t slot[NS][BMW];
    Handoff handoff[NS];
    uint32_t state[NS];
    uint32_t done;
    uint32_t stage[NE][CH];
};

// Ordering of the LDS handoffs between the waves of a block (LDS only).
__device__ __forceinline__ void lds_release() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local"); }
__device__ __forceinline__ void lds_acquire() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local"); }

// A free slot taken for filling, or -1 when every slot is busy.
__device__ __forceinline__ int take_slot(uint32_t *state) {
    int s = -1;
    if (lane() == 0) {
#pragma unroll
        for (uint32_t i = 0; i < NS; i++) {
            uint32_t expect = SLOT_FREE;
            if (__hip_atomic_compare_exchange_strong(&state[i], &expect, SLOT_FILLING, __ATOMIC_ACQUIRE,
                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) {
                s = (int)i;
                break;
            }
        }
    }
    return (int)rdl((uint32_t)s, 0);
}

// An emitter wave of a block (with output): goes round the slots, takes a
// ready one (ready -> emitting), and when the look-back finds the segment's
// offset without waiting writes its matches and, for a pair's last segment,
// the pair's count, then frees the slot; otherwise it makes the slot ready
// again for the next pass. A slot that waits longer than the look-back limit
// is given up (the call reports it). Ends when every builder of the block has
// finished and no slot is ready. Its stores never hold up a builder's loads
// (each wave counts its own memory operations).
__device__ void emitter(BlockLds<true> &sh, uint32_t *stage, unsigned long long *status, uint32_t epoch,
                        uint32_t *work, uint32_t *counts, uint32_t *out) {
    const uint32_t l = lane();
    while (true) {
        bool ready_any = false, progress = false;
#pragma unroll
        for (uint32_t i = 0; i < NS; i++) {
            if (rdl(__hip_atomic_load(&sh.state[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP), 0) != SLOT_READY)
                continue;
            ready_any = true;
            uint32_t took = 0;
            if (l == 0) {
                uint32_t expect = SLOT_READY;
                took = __hip_atomic_compare_exchange_strong(&sh.state[i], &expect, SLOT_EMITTING, __ATOMIC_ACQUIRE,
                                                            __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            if (!rdl(took, 0)) continue;
            Handoff &h = sh.handoff[i];
            const uint32_t g = __builtin_amdgcn_readfirstlane(h.g), sb = __builtin_amdgcn_readfirstlane(h.sb);
            const uint32_t c = __builtin_amdgcn_readfirstlane(h.c);
            uint32_t prefix = 0;
            uint32_t next = SLOT_FREE;
            if (!try_resolve(status, g, sb, epoch, c, prefix)) {
                const unsigned long long now = __builtin_amdgcn_s_memrealtime(), since = h.since;
                if (!since) {
                    if (l == 0) h.since = now;
                    next = SLOT_READY;
                } else if (now - since <= WAIT_TICKS) {
                    next = SLOT_READY;
                } else if (l == 0) {
                    atomicOr(&work[ISX_FLAGS], ISX_GAVE_UP);
                }
            } else {
                const uint32_t room = __builtin_amdgcn_readfirstlane(h.room), nw = __builtin_amdgcn_readfirstlane(h.nw);
                const uint32_t amin = __builtin_amdgcn_readfirstlane(h.amin), p = __builtin_amdgcn_readfirstlane(h.p);
                const uint32_t last = __builtin_amdgcn_readfirstlane(h.last);
                const uint64_t off = h.off;
                if (c) dense_emit(sh.slot[i], nw, amin, stage, out + off + prefix, prefix < room ? room - prefix : 0u);
                if (last && l == 0) counts[p] = prefix + c;
                progress = true;
            }
            lds_release();
            if (l == 0) __hip_atomic_store(&sh.state[i], next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        if (progress) continue;
        if (!ready_any &&
            rdl(__hip_atomic_load(&sh.done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP), 0) == NB) {
            lds_acquire();
            bool busy = false;
            for (uint32_t i = 0; i < NS; i++)
                busy |= __hip_atomic_load(&sh.state[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == SLOT_READY;
            if (!__ballot(busy)) break;
            continue;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// Claims: the segments are dealt to ncls <= ISX_CLASSES classes in runs of
// RUN consecutive segments (run r to class r % ncls), and each class's
// counter hands out its segments in order; a builder wave claims from the
// class of its block (blocks are dispatched to the XCDs round-robin, so each
// XCD claims mostly from its own counter, and neighbouring segments of a run,
// whose merge-path cut searches read each other's boundary lines, share an
// L2) and, once that class is exhausted, from the others in turn. One
// counter for all segments serialised the claims (about 70 per microsecond
// on one address).
// With output (OUT) a block has NB builder waves and NE emitter waves. A
// builder hands a dense segment's common bits to the emitters through a free
// slot and goes on to its next segment, so its loads are never held up by
// stores of its own; with no free slot, and for sparse and skew segments, it
// finds the offset and writes the matches itself.
// A wave publishes the count of every segment it builds before it waits on
// any other, waits only on lower segments of the same pair, and a builder
// never waits for the emitter. Every wait ends: a claimed segment is
// published by the builder that claimed it once it is built; a wait on an
// unclaimed lower segment t of class j by a wave whose own segment is of
// class i means counter_j <= t < own < counter_i, so such waits form no
// cycle, and the builders of the class with the lowest counter (every class
// has builders: ncls <= the grid's blocks) wait only on claimed segments.
template <bool OUT>
__global__ __launch_bounds__(W *(OUT ? NB + NE : WPB_COUNT)) void fm_isx_kernel(
    const uint32_t *__restrict__ data, const uint64_t *a_off, const uint32_t *a_len, const uint64_t *b_off,
    const uint32_t *b_len, uint32_t npairs, uint32_t *counts, const uint64_t *out_off, uint32_t *out, uint32_t *work,
    uint32_t cap, uint32_t ncls) {
    __shared__ BlockLds<OUT> sh;
    const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x / W), l = lane();
    const uint32_t total = __builtin_amdgcn_readfirstlane(ld_rlx(&work[ISX_TOTAL]));
    const uint32_t epoch = __builtin_amdgcn_readfirstlane(ld_rlx(&work[ISX_EPOCH]));
    const uint32_t *const seg_begin = work + isx_table_word(cap);
    unsigned long long *const status = reinterpret_cast<unsigned long long *>(work + ISX_HDR);
    if constexpr (OUT) {
        if (threadIdx.x < NS) sh.state[threadIdx.x] = SLOT_FREE;
        if (threadIdx.x == 0) sh.done = 0;
        __syncthreads();
        if (wv >= NB) {
            emitter(sh, sh.stage[wv - NB], status, epoch, work, counts, out);
            return;
        }
    }
    uint32_t *const bmA = sh.builder[wv], *const bmB = bmA + BMW, *const cs = bmA + 2 * BMW;
    uint32_t cls = blockIdx.x % ncls, tried = 0;
    while (true) {
        uint32_t g = total;
        while (tried < ncls) {
            // segment of the class's k-th claim
            auto seg_of = [&](uint32_t k) { return ((unsigned long long)(k / RUN) * ncls + cls) * RUN + k % RUN; };
            uint32_t k = 0;
            if (l == 0) {
                uint32_t *const ctr = &work[ISX_CLAIM + 32 * cls];
                // after the home class, a class is read before it is claimed
                // from: claims on exhausted classes by every wave at the end
                // would queue on the counters' addresses
                bool spent = false;
                if (tried) {
                    k = ld_rlx(ctr);
                    spent = seg_of(k) >= total;
                }
                if (!spent) k = __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            k = rdl(k, 0);
            const unsigned long long gg = seg_of(k);
            if (gg < total) {
                g = __builtin_amdgcn_readfirstlane((uint32_t)gg);
                break;
            }
            cls = cls + 1 == ncls ? 0u : cls + 1;  // this class is exhausted
            tried++;
        }
        if (g >= total) break;
        const uint32_t p = __builtin_amdgcn_readfirstlane(find_pair(seg_begin, npairs, g));
        const uint32_t sb = ldc(seg_begin + p), s = g - sb, last = g + 1 == ldc(seg_begin + p + 1);
        const Pair q = pair_of(data, a_off, a_len, b_off, b_len, p);
        const uint64_t ooff = OUT ? ldc(out_off + p) : 0ull;
        if ((uint64_t)q.nb >= (uint64_t)ISX_SKEW * q.na) {
            // skew segment: CH elements of A, found in B and kept in registers
            const uint32_t first = s * CH;
            uint32_t a[K];
            bool found[K];
            const uint32_t c = skew_find(q.A + first, min(CH, q.na - first), q.B, q.nb, a, found);
            if (!OUT) {
                if (l == 0 && c) __hip_atomic_fetch_add(&counts[p], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                continue;
            }
            publish(status, g, sb, epoch, c);
            const uint32_t prefix = resolve(status, g, sb, epoch, c, work);
            uint32_t *const o = out + ooff + prefix;
            const uint32_t lim = prefix < q.na ? q.na - prefix : 0u;
            uint32_t w = 0;
#pragma unroll
            for (uint32_t k = 0; k < K; k++) w = emit<true>(found[k], a[k], w, o, lim);
            if (last && l == 0) counts[p] = prefix + c;
            continue;
        }
        // merge segment: the merge-path splits of its first and last
        // diagonals (lanes 0-31 and 32-63)
        const uint64_t d0 = (uint64_t)s * ISX_SEG, d1 = min(d0 + ISX_SEG, (uint64_t)q.na + q.nb);
        const uint64_t d = l < 32 ? d0 : d1;
        const uint32_t *const A = q.A, *const B = q.B;
        const uint32_t r = first_false2(d > q.nb ? d - q.nb : 0, min(d, (uint64_t)q.na),
                                        [&](uint64_t i) { return A[i] <= B[d - 1 - i]; });
        const uint32_t i0 = rdl(r, 0), i1 = rdl(r, 32);
        const uint32_t j0 = (uint32_t)(d0 - i0), j1 = (uint32_t)(d1 - i1);
        const uint32_t jx = min(j1 + 1, q.nb);
        const uint32_t la = i1 - i0, lb = jx - j0;
        const uint32_t amin = la ? ldc(A + i0) : 0u, span = la ? ldc(A + i1 - 1) - amin : 0u;
        if (span < BMSPAN) {
            dense_build(A + i0, la, B + j0, lb, amin, span, bmA, bmB);
            const uint32_t nw = la ? (span >> 5) + 1 : 0u;
            if constexpr (!OUT) {
                const uint32_t c = common_bits<false>(bmA, bmB, nw, nullptr);
                if (l == 0 && c) __hip_atomic_fetch_add(&counts[p], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                const int slot = take_slot(sh.state);
                const uint32_t c = common_bits<true>(bmA, bmB, nw, slot >= 0 ? sh.slot[slot] : bmA);
                publish(status, g, sb, epoch, c);
                if (!c && !last) {
                    // nothing to write, and no pair count: the published count is all later segments need
                    if (slot >= 0 && l == 0)
                        __hip_atomic_store(&sh.state[slot], SLOT_FREE, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                } else if (slot >= 0) {
                    if (l == 0) sh.handoff[slot] = Handoff{g, sb, p, nw, amin, c, last, q.na, ooff, 0ull};
                    lds_release();
                    if (l == 0)
                        __hip_atomic_store(&sh.state[slot], SLOT_READY, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                } else {
                    const uint32_t prefix = resolve(status, g, sb, epoch, c, work);
                    if (c) dense_emit(bmA, nw, amin, cs, out + ooff + prefix, prefix < q.na ? q.na - prefix : 0u);
                    if (last && l == 0) counts[p] = prefix + c;
                }
            }
            continue;
        }
        // sparse segment: counted first; its matches are found again and
        // written once its offset is known
        const uint32_t c = sparse_segment<false>(A + i0, la, B + j0, lb, cs, nullptr, 0);
        if (!OUT) {
            if (l == 0 && c) __hip_atomic_fetch_add(&counts[p], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            continue;
        }
        publish(status, g, sb, epoch, c);
        const uint32_t prefix = resolve(status, g, sb, epoch, c, work);
        if (c) sparse_segment<true>(A + i0, la, B + j0, lb, cs, out + ooff + prefix, prefix < q.na ? q.na - prefix : 0u);
        if (last && l == 0) counts[p] = prefix + c;
    }
    if constexpr (OUT) {
        lds_release();
        if (l == 0) __hip_atomic_fetch_add(&sh.done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
}

template <bool OUT>
int blocks_per_cu() {
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, fm_isx_kernel<OUT>, W * (OUT ? NB + NE : WPB_COUNT), 0) !=
            hipSuccess ||
        b <= 0)
        b = 1;
    return b;
}

}  // namespace

void isx_launch(hipStream_t stream, uint32_t cus, uint32_t *work, uint32_t cap, const uint32_t *data,
                const uint64_t *a_off, const uint32_t *a_len, const uint64_t *b_off, const uint32_t *b_len,
                uint32_t npairs, uint32_t *counts, const uint64_t *out_off, uint32_t *out) {
    hipLaunchKernelGGL(fm_isx_plan_kernel, dim3(1), dim3(1024), 0, stream, a_len, b_len, npairs, counts, work, cap);
    const uint32_t grid = std::max<uint32_t>(1u, (uint32_t)isx_blocks_per_cu(out_off != nullptr) * cus);
    const uint32_t ncls = std::min<uint32_t>(ISX_CLASSES, grid);
    if (out_off)
        hipLaunchKernelGGL(fm_isx_kernel<true>, dim3(grid), dim3(W * (NB + NE)), 0, stream, data, a_off, a_len, b_off,
                           b_len, npairs, counts, out_off, out, work, cap, ncls);
    else
        hipLaunchKernelGGL(fm_isx_kernel<false>, dim3(grid), dim3(W * WPB_COUNT), 0, stream, data, a_off, a_len,
                           b_off, b_len, npairs, counts, out_off, out, work, cap, ncls);
}

int isx_blocks_per_cu(bool out) {
    static const int with_out = blocks_per_cu<true>(), count_only = std::min(COUNT_BLOCKS_PER_CU, blocks_per_cu<false>());
    return out ? with_out : count_only;
}

}  // namespace fm
