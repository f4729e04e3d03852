// ForkMatch MI355X engine: the persistent subgraph-matching kernel, its
// device work queues, the bank pool, the paged output table, the wave64
// intersection microkernel, and the device-side C ABI.
//
// Reference counterparts (src/ = the reference tree):
//  - initPosting/isoInit/isoLoop/isoMatch (src/smsm.cu:395-556,784-942):
//    one persistent launch; each wave64 repeatedly takes a task (a range of
//    partial embeddings of one query at one depth) and extends it one depth.
//  - the dynamic-parallelism fork chain (src/smsm.cu:559-687,947-1139): a
//    full output bank is not forked as a child grid; its rows are pushed as
//    tasks onto device queues ordered by remaining depth, so deeper work
//    runs first and memory stays bounded.
//  - MemoryPool/forest (src/tbs.cuh:10-136): banks are carved by a device
//    bump allocator from one arena and recycled through a device free list
//    with reference counts. Partial embeddings are stored as full rows, not
//    as (mapping, prev) chains, so a consumer reads a parent with one
//    coalesced load instead of a dependent pointer walk.
//  - isoGen/isoCompare/isInj (src/smsm.cu:55-140,187-227,297-345): the
//    candidates of a batch of up to 64 parents are flattened with a wave
//    prefix sum over the shortest requirement row of each parent, so every
//    lane checks one candidate per step whatever the row lengths are;
//    membership in the other rows is a binary search, injectivity a compare
//    against the parent row in LDS.
//  - isoWrite (src/smsm.cu:704-779): final-depth rows are written straight
//    into lazily allocated output pages, columns in query-vertex order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <thread>
#include <vector>

#include "fm_internal.hpp"
#include "isect.hpp"

namespace fm {

constexpr uint32_t WAVE = 64;
constexpr uint32_t MAXN = FM_MAX_QUERY_VERTICES;
constexpr uint32_t MAXR = FM_MAX_REQS_PER_DEPTH;
// Smallest bank (8-byte units): a fresh bank holds one paired candidate
// step's children, 2 x WAVE rows of at most MAXN words.
constexpr uint32_t MIN_BANK_SIZE = WAVE * MAXN;
constexpr uint32_t NLEVELS = MAXN;
constexpr uint32_t LABEL_ABSENT = 0xFFFFFFu;
constexpr uint32_t ROOT_BANK = 0xFFFFFFFFu;
constexpr uint32_t NO_BANK = 0xFFFFFFFFu;
constexpr uint32_t ALL_VERTICES = 0xFFFFFFFFu;
constexpr uint32_t NO_PAGE = 0xFFFFFFFFu;
// serving: a table page whose allocation failed (its rows are not stored)
constexpr uint32_t FAILED_PAGE = 0xFFFFFFFEu;

// Packed requirement: bits 0-5 anchor depth (== own depth: self-loop),
// bit 7 direction (0 out-row, 1 in-row), bits 8-31 dense label id.
__host__ __device__ inline uint32_t req_anchor(uint32_t r) { return r & 63u; }
__host__ __device__ inline uint32_t req_dir(uint32_t r) { return (r >> 7) & 1u; }
__host__ __device__ inline uint32_t req_label(uint32_t r) { return r >> 8; }

constexpr uint32_t MAXAUT = FM_MAX_AUTOMORPHISMS;
constexpr uint32_t MAXSC = 64;  // symmetry-breaking constraints per plan

struct DPlan {
    uint32_t n;
    uint32_t root_base;  // ALL_VERTICES or offset into the combined root list
    uint32_t nroots;
    uint32_t naut;       // > 1: symmetry breaking, every embedding found is expanded naut ways
    uint32_t order[MAXN];
    uint32_t req_begin[MAXN + 1];
    uint32_t nself[MAXN];  // number of self-loop requirements per depth
    uint32_t req[FM_MAX_REQS_TOTAL];  // per depth: non-self first, then self
    // ordering constraints of each depth: anchor depth, bit 7 set when the
    // candidate must be greater than the anchor's vertex (else smaller)
    uint32_t sc_begin[MAXN + 1];
    uint32_t sc[MAXSC];
    uint8_t autcol[MAXAUT][MAXN];  // expansion k: column u takes the vertex of depth autcol[k][u]
};

struct DQuery {
    unsigned long long count;
    unsigned long long rows;       // table rows stored
    unsigned long long row_words;  // instrumented
    unsigned long long lookups;
    unsigned long long candidates;  // instrumented
    unsigned long long probes;      // instrumented
    unsigned long long partials[MAXN];
    unsigned long long max_rows;
    unsigned long long max_count;
    unsigned long long max_partials;
    unsigned long long stored;  // intermediate partial embeddings written
    unsigned long long spare;  // partly filled page handed between waves: (id + 1) << 32 | fill
    long long qpending;        // serving: this query's unfinished tasks (completion)
    // serving tables are dense: row r sits in page r / page_rows of the slot
    unsigned long long cursor;    // rows claimed
    unsigned long long trunc_at;  // rows from here on are not stored (budget or arena exhausted)
    uint32_t status;
    uint32_t flags;
    uint32_t page_rows;
    uint32_t max_pages;
    uint32_t npages;  // pages claimed
    uint32_t root_chunk_begin;
    uint32_t chunk_roots;  // roots per root claim of this query
    uint32_t count_only;   // 1: counted without rows in a table-mode launch
    uint32_t pad;
};

enum StatIdx {
    ST_TASKS = 0,
    ST_ROOTS,
    ST_BANKS_FRESH,
    ST_BANKS_REUSED,
    ST_ALLOC_FAIL,
    ST_PAGES,
    ST_REQUEUE,
    ST_SPLIT,
    ST_BUSY,   // wall-clock ticks waves spent inside tasks and admissions
    ST_ALIVE,  // wall-clock ticks waves were resident
    ST_STALE,  // serving: tasks of already completed queries, dropped
    ST_N
};

// Control block. The first part is cleared before every launch; the queue
// and pool positions persist across launches (both are empty at the end of
// a clean launch) and are re-initialised after an aborted one. Words that
// different waves hammer sit on separate 128-byte lines.
struct LevelQ {
    alignas(128) uint32_t head;
    alignas(128) uint32_t tail;
    alignas(128) int32_t avail;
};

// Idle-wave count in shards on separate lines: waves going idle and busy
// update one shard each, readers sum them.

constexpr unsigned long long SRV_ACTIVE_ONE = 1ull << 32;
constexpr unsigned long long SRV_CLOSING = 1ull << 63;
__host__ __device__ inline uint32_t srv_head(unsigned long long st) { return (uint32_t)st; }
__host__ __device__ inline uint32_t srv_active(unsigned long long st) {
    return (uint32_t)((st & ~SRV_CLOSING) >> 32);
}

constexpr uint32_t STAT_SHARDS = 64;
constexpr uint32_t STAT_STRIDE = 16;  // u64 per shard: one 128-byte line
static_assert(ST_N <= STAT_STRIDE, "counters fit a shard");

struct Ctl {
    alignas(128) long long pending;  // tasks pushed or root chunks not yet finished
    // Batch launches: set by the wave whose task end takes pending to zero
    // (pending never rises again from zero there), or by the host for a
    // launch without root chunks. Waiting waves poll this quiet word rather
    // than pending, which every task end updates.
    alignas(128) uint32_t done;
    alignas(128) unsigned long long root_cursor;
    // Waves that found no work. It changes only when a wave runs out of work
    // or finds some again, so one word takes the atomics; busy waves read it
    // at their work-splitting checks.
    alignas(128) int32_t idle_waves;
    alignas(128) uint32_t abort;
    uint32_t error;
    unsigned long long t0;  // wall clock of the first wave (deadline origin)
    unsigned long long timeout_ticks;
    uint32_t diag[12];  // FM_ERR_INTERNAL: the task; FM_ERR_TIMEOUT: where the launch ran out of time
    alignas(128) uint32_t page_bump;
    // per-launch counters, summed by the host over the shards: each wave adds
    // its counters into shard (workgroup id % STAT_SHARDS) at exit, so
    // thousands of waves leaving together do not queue on one cache line
    alignas(128) unsigned long long stats[STAT_SHARDS][STAT_STRIDE];
    // persistent
    alignas(128) LevelQ lq[NLEVELS];
    alignas(128) int32_t queued;  // tasks in the queues, all levels (sum of lq[].avail)
    alignas(128) uint32_t f_head;
    alignas(128) uint32_t f_tail;
    alignas(128) int32_t f_avail;
    alignas(128) uint32_t bank_bump;
    // serving (persistent across serving launches)
    // Serving state in one word, so that a claim, the count of queries in
    // flight and the decision to end the launch change together: bits 0-31
    // the next submission to admit, bits 32-62 the admitted queries not yet
    // completed, bit 63 set once the launch is closing (no more claims; all
    // waves leave). The host clears bit 63 before it relaunches.
    alignas(128) unsigned long long srv_state;
    alignas(128) int32_t live_banks;     // serving: banks out of the pool
    alignas(128) uint32_t srv_page_bump;
    alignas(128) uint32_t pf_head;       // recycled output pages
    alignas(128) uint32_t pf_tail;
    alignas(128) int32_t pf_avail;
};
constexpr size_t CTL_RESET_BYTES = offsetof(Ctl, lq);

struct alignas(32) Slot {
    uint32_t seq;
    uint32_t q;
    uint32_t depth;
    uint32_t bank;
    uint32_t begin;  // word offset in the bank, or first root index
    uint32_t count;
    uint32_t skip;      // driver-row offset where the first parent starts
    uint32_t last_end;  // driver-row offset where the last parent stops
};

struct alignas(8) FSlot {
    uint32_t seq;
    uint32_t id;
};

// Serving: host-pinned, coherent memory shared with the host.
constexpr uint32_t SUB_CAP = 4096;
struct SrvSub {
    uint32_t slot;
    uint32_t ticket;
    uint32_t pad[2];
};
struct SrvQInit {
    unsigned long long max_rows;
    uint32_t page_rows;
    uint32_t max_pages;
    uint32_t ticket;
    uint32_t pad[3];
};
struct SrvDone {
    uint32_t count_lo, count_hi, rows_lo, rows_hi;
    uint32_t status, flags, npages, ticket;  // ticket is written last
};
struct SrvShared {
    uint32_t sub_tail;    // host: submissions written
    uint32_t shutdown;    // host: stop when idle
    uint32_t admitted;    // device: submissions admitted
    uint32_t live_banks;  // device: banks out of the pool (admission budget)
    uint32_t abort;       // host: end the launch now (shutdown that did not drain)
    uint32_t pad[27];
    SrvSub sub[SUB_CAP];
};

struct DGraph {
    uint32_t V;
    uint32_t L;
    const uint32_t *ptr[2];  // row pointers into col (in-rows offset by E)
    const uint32_t *col;
    const uint32_t *skip;  // col[b * 32] for every block b (membership probes)
    const uint32_t *roots;
};

struct KParams {
    DGraph g;
    const DPlan *plans;
    DQuery *qs;
    Ctl *ctl;
    Slot *ring;
    FSlot *free_ring;
    int32_t *bank_ref;
    uint32_t *banks;
    uint32_t *pages;
    uint32_t *page_owner;  // query of each claimed output page
    uint32_t *page_fill;   // rows held by each output page
    unsigned long long timeout_ticks;
    unsigned long long starve_ticks;
    unsigned long long bank_words;
    unsigned long long page_words;
    uint32_t nq;
    uint32_t total_root_chunks;
    uint32_t root_chunk;
    uint32_t task_rows;
    uint32_t tile_parents;  // parents staged in LDS per batch (1..WAVE)
    uint32_t table;
    uint32_t instrument;
    uint32_t qcap;
    uint32_t fcap;
    uint32_t nbanks;
    uint32_t npages;
    uint32_t nlevels;
    uint32_t serving;
    // serving
    SrvShared *srv;
    const DPlan *srv_plans;    // pinned: plan of each slot
    const SrvQInit *srv_qinit; // pinned
    SrvDone *srv_done;         // pinned
    DPlan *plans_rw;           // device copy of the slot plans (same as plans)
    uint32_t *slot_pages;      // [max_slots * max_slot_pages]
    FSlot *pfree_ring;
    uint32_t pfcap;
    uint32_t max_slots;
    uint32_t max_slot_pages;
    // Waves with workgroup id >= active_waves (when nonzero) start dormant
    // and join the launch once it has run for join_ticks, unless it has
    // ended by then.
    uint32_t active_waves;
    unsigned long long idle_ticks;
    unsigned long long slice_ticks;  // serving: the launch retires after this long (0 = never)
    unsigned long long join_ticks;
};

// ---------------------------------------------------------------------------
// Device helpers.

__device__ __forceinline__ uint32_t ld_acq(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t ld_rlx(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int32_t ld_rlx(const int32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ long long ld_rlx(const long long *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long ld_rlx(const unsigned long long *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rel(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rlx(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ T atom_add(T *p, T v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool cas_u32(uint32_t *p, uint32_t expect, uint32_t desired) {
    return __hip_atomic_compare_exchange_strong(p, &expect, desired, __ATOMIC_ACQ_REL, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void backoff() { __builtin_amdgcn_s_sleep(2); }

struct Ctl;
__device__ __forceinline__ bool spin_ok(const Ctl *c, uint32_t &iter, uint32_t site, uint32_t a, uint32_t b);

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Lane src's value on every lane (src is the same on every lane): a scalar
// read of the lane's register, no LDS round trip.
__device__ __forceinline__ uint32_t bcast(uint32_t v, int src) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, src);
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int o = 1; o < (int)WAVE; o <<= 1) {
        uint32_t t = (uint32_t)__shfl_up((int)v, o, WAVE);
        if (lane >= (uint32_t)o) v += t;
    }
    return v;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
    uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
    // sum the halves separately with carries folded through 64-bit adds
    unsigned long long acc = v;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        lo = (uint32_t)acc;
        hi = (uint32_t)(acc >> 32);
        uint32_t olo = (uint32_t)__shfl_xor((int)lo, o, WAVE);
        uint32_t ohi = (uint32_t)__shfl_xor((int)hi, o, WAVE);
        acc += ((unsigned long long)ohi << 32) | olo;
    }
    return acc;
}

// COL_PAD, SKIP_SHIFT: the device column array's padding and skip index
// (fm_internal.hpp).
constexpr uint32_t SKIP_BLOCK = 1u << SKIP_SHIFT;

// The five 16-byte words of a[w0, w0 + 20) for w0 = lo rounded down to 4
// words: a window that holds any 17 consecutive positions from lo. All five
// loads are issued before any is used (one round trip).
struct Window20 {
    uint4 q[5];
};
__device__ __forceinline__ Window20 load_window(const uint32_t *a, uint32_t lo) {
    const uint4 *wp = reinterpret_cast<const uint4 *>(a + (lo & ~3u));
    Window20 w;
#pragma unroll
    for (uint32_t i = 0; i < 5; i++) w.q[i] = wp[i];
    return w;
}
__device__ __forceinline__ uint32_t window_word(const Window20 &w, uint32_t j) {
    const uint4 &q = w.q[j >> 2];
    return (j & 3) == 0 ? q.x : (j & 3) == 1 ? q.y : (j & 3) == 2 ? q.z : q.w;
}
// Bits j of the window's words equal to x / at most x: one compare and one
// select per word, no per-word range test (the caller masks the range once).
// The bits of one 16-byte load are combined before the next load's, so few
// intermediate values are live at once.
__device__ __forceinline__ uint32_t window_eq_bits(const Window20 &w, uint32_t x) {
    uint32_t m = 0;
#pragma unroll
    for (uint32_t i = 0; i < 5; i++) {
        const uint4 q = w.q[i];
        const uint32_t g = (q.x == x ? 1u : 0u) | (q.y == x ? 2u : 0u) | (q.z == x ? 4u : 0u) | (q.w == x ? 8u : 0u);
        m |= g << (4 * i);
    }
    return m;
}
__device__ __forceinline__ uint32_t window_le_bits(const Window20 &w, uint32_t x) {
    uint32_t m = 0;
#pragma unroll
    for (uint32_t i = 0; i < 5; i++) {
        const uint4 q = w.q[i];
        const uint32_t g = (q.x <= x ? 1u : 0u) | (q.y <= x ? 2u : 0u) | (q.z <= x ? 4u : 0u) | (q.w <= x ? 8u : 0u);
        m |= g << (4 * i);
    }
    return m;
}
// Window positions [r, r + n) for n <= 16, r <= 3.
__device__ __forceinline__ uint32_t range_bits(uint32_t r, uint32_t n) { return ((1u << n) - 1u) << r; }

// The number of skip entries of a row of more than 16 words: the blocks
// that start inside it, (st >> 4) + 1 .. (st + len - 1) >> 4.
__device__ __forceinline__ uint32_t skip_entries(uint32_t st, uint32_t len) {
    return ((st + len - 1) >> SKIP_SHIFT) - (st >> SKIP_SHIFT);
}

// Membership of x in the sorted, strictly increasing row col[st, st + len).
// A row of at most 16 words is one window. A longer row first counts, in one
// window of the skip index, the blocks starting inside the row whose first
// word is <= x (binary steps on the skip index first narrow rows of more
// than 17 blocks): the last such block, or the row's first block, is the
// only one that can hold x, and a second window reads it. So a probe costs
// two round trips up to 272 words, plus one per doubling past that on the
// small skip index, instead of a chain of ~log2(len) dependent loads.
__device__ __forceinline__ bool row_contains(const uint32_t *col, const uint32_t *skip, uint32_t st, uint32_t len,
                                             uint32_t x) {
    if (len == 0) return false;
    uint32_t lo = st, end = st + len;
    if (len > SKIP_BLOCK) {
        const uint32_t b0 = st >> SKIP_SHIFT, b1 = (end - 1) >> SKIP_SHIFT;
        // blocks b0+1 .. b1 start inside the row; those below blo hold a
        // first word <= x, those at or past blo + bn one > x
        uint32_t blo = b0 + 1, bn = b1 - b0;
        while (bn > 16) {
            const uint32_t half = bn >> 1;
            if (skip[blo + half] <= x) {
                blo += half + 1;
                bn -= half + 1;
            } else {
                bn = half;
            }
        }
        const Window20 w = load_window(skip, blo);
        const uint32_t c = __popc(window_le_bits(w, x) & range_bits(blo & 3u, bn));
        const uint32_t bucket = blo - 1 + c;
        lo = max(st, bucket << SKIP_SHIFT);
        end = min(end, (bucket + 1) << SKIP_SHIFT);
    }
    // x, if present, lies in [lo, end), at most 16 positions
    const Window20 w = load_window(col, lo);
    return (window_eq_bits(w, x) & range_bits(lo & 3u, end - lo)) != 0;
}

// Membership of xa in row a and of xb in row b by the steps of
// row_contains, with both rows' loads in flight together: the candidate loop
// probes two candidates per lane, so each step waits once per probe level
// instead of twice. A row of length 0 holds nothing; its loads read the
// first words of the arrays, which always exist.
__device__ __forceinline__ void row_contains2(const uint32_t *col, const uint32_t *skip, uint32_t sta, uint32_t lena,
                                              uint32_t xa, uint32_t stb, uint32_t lenb, uint32_t xb, bool &ha,
                                              bool &hb) {
    uint32_t loa = sta, enda = sta + lena, lob = stb, endb = stb + lenb;
    const bool sa = lena > SKIP_BLOCK, sb = lenb > SKIP_BLOCK;
    if (sa || sb) {
        uint32_t bloa = 0, bna = 0, blob = 0, bnb = 0;
        if (sa) {
            bloa = (sta >> SKIP_SHIFT) + 1;
            bna = skip_entries(sta, lena);
        }
        if (sb) {
            blob = (stb >> SKIP_SHIFT) + 1;
            bnb = skip_entries(stb, lenb);
        }
        while (bna > 16 || bnb > 16) {
            const uint32_t hfa = bna >> 1, hfb = bnb >> 1;
            const uint32_t va = skip[bna > 16 ? bloa + hfa : 0u];
            const uint32_t vb = skip[bnb > 16 ? blob + hfb : 0u];
            if (bna > 16) {
                if (va <= xa) {
                    bloa += hfa + 1;
                    bna -= hfa + 1;
                } else {
                    bna = hfa;
                }
            }
            if (bnb > 16) {
                if (vb <= xb) {
                    blob += hfb + 1;
                    bnb -= hfb + 1;
                } else {
                    bnb = hfb;
                }
            }
        }
        const Window20 wa = load_window(skip, bloa), wb = load_window(skip, blob);
        const uint32_t ca = __popc(window_le_bits(wa, xa) & range_bits(bloa & 3u, bna));
        const uint32_t cb = __popc(window_le_bits(wb, xb) & range_bits(blob & 3u, bnb));
        if (sa) {
            const uint32_t bucket = bloa - 1 + ca;
            loa = max(sta, bucket << SKIP_SHIFT);
            enda = min(enda, (bucket + 1) << SKIP_SHIFT);
        }
        if (sb) {
            const uint32_t bucket = blob - 1 + cb;
            lob = max(stb, bucket << SKIP_SHIFT);
            endb = min(endb, (bucket + 1) << SKIP_SHIFT);
        }
    }
    const Window20 wa = load_window(col, loa), wb = load_window(col, lob);
    ha = (window_eq_bits(wa, xa) & range_bits(loa & 3u, enda - loa)) != 0;
    hb = (window_eq_bits(wb, xb) & range_bits(lob & 3u, endb - lob)) != 0;
}

__device__ __forceinline__ void row_of(const DGraph &g, uint32_t req, uint32_t v, uint32_t &st, uint32_t &len) {
    uint32_t lab = req_label(req);
    if (lab == LABEL_ABSENT) {
        st = 0;
        len = 0;
        return;
    }
    const uint32_t *p = g.ptr[req_dir(req)] + ((size_t)lab * g.V + v);
    st = p[0];
    len = p[1] - st;
}

// Every spin-wait calls this between polls: it sleeps briefly and, every
// 256 polls, gives up (returns false) once the launch is aborting or past
// its deadline, so no wait can outlive the launch.
// Records why the launch timed out (first caller wins): diag[0] = site,
// diag[1..3] = site-specific values.
__device__ __forceinline__ void timeout_at(const Ctl *c, uint32_t site, uint32_t a, uint32_t b, uint32_t d) {
    uint32_t zero = 0;
    Ctl *w = const_cast<Ctl *>(c);
    if (__hip_atomic_compare_exchange_strong(&w->error, &zero, (uint32_t)FM_ERR_TIMEOUT, __ATOMIC_RELAXED,
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        w->diag[0] = site;
        w->diag[1] = a;
        w->diag[2] = b;
        w->diag[3] = d;
    }
    __hip_atomic_store(&w->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool spin_ok(const Ctl *c, uint32_t &iter, uint32_t site, uint32_t a, uint32_t b) {
    backoff();
    if ((++iter & 255u) != 0) return true;
    if (__hip_atomic_load(&c->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
    unsigned long long t0 = __hip_atomic_load(&c->t0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned long long lim = __hip_atomic_load(&c->timeout_ticks, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (t0 && wall_clock64() - t0 > lim) {
        timeout_at(c, site, a, b, 0);
        return false;
    }
    return true;
}

// Waits until *p == want, polling with relaxed loads (no cache invalidation
// per poll) and acquiring once at the end. false: the launch is aborting.
// `site` names the wait in a timeout report.
__device__ __forceinline__ bool wait_eq(const uint32_t *p, uint32_t want, const Ctl *c, uint32_t site) {
    uint32_t it = 0;
    uint32_t v;
    while ((v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) != want)
        if (!spin_ok(c, it, site, v, want)) return false;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return true;
}

__device__ __forceinline__ void raise_error(const KParams &P, uint32_t code) {
    uint32_t zero = 0;
    __hip_atomic_compare_exchange_strong(&P.ctl->error, &zero, code, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
    st_rlx(&P.ctl->abort, 1u);
}

__device__ __forceinline__ void query_fail(const KParams &P, uint32_t q, uint32_t code) {
    uint32_t zero = 0;
    __hip_atomic_compare_exchange_strong(&P.qs[q].status, &zero, code, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
}

// ---- bank pool (lane 0 only) ----------------------------------------------

__device__ __forceinline__ uint32_t bank_alloc(const KParams &P, unsigned long long *stats) {
    Ctl *c = P.ctl;
    if (ld_rlx(&c->f_avail) > 0) {
        int32_t a = atom_add(&c->f_avail, -1);
        if (a > 0) {
            uint32_t pos = atom_add(&c->f_head, 1u);
            FSlot *s = &P.free_ring[pos & (P.fcap - 1)];
            if (!wait_eq(&s->seq, pos + 1, c, 11)) return NO_BANK;
            uint32_t id = s->id;
            st_rel(&s->seq, pos + P.fcap);
            __hip_atomic_store(&P.bank_ref[id], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (P.serving) atom_add(&c->live_banks, 1);
            stats[ST_BANKS_REUSED]++;
            return id;
        }
        atom_add(&c->f_avail, 1);
    }
    if (ld_rlx(&c->bank_bump) < P.nbanks) {
        uint32_t id = atom_add(&c->bank_bump, 1u);
        if (id < P.nbanks) {
            __hip_atomic_store(&P.bank_ref[id], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (P.serving) atom_add(&c->live_banks, 1);
            stats[ST_BANKS_FRESH]++;
            return id;
        }
    }
    return NO_BANK;
}

__device__ __forceinline__ void bank_release(const KParams &P, uint32_t id) {
    if (atom_add(&P.bank_ref[id], -1) == 1) {
        Ctl *c = P.ctl;
        uint32_t pos = atom_add(&c->f_tail, 1u);
        FSlot *s = &P.free_ring[pos & (P.fcap - 1)];
        if (!wait_eq(&s->seq, pos, c, 12)) return;
        s->id = id;
        st_rel(&s->seq, pos + 1);
        atom_add(&c->f_avail, 1);
        if (P.serving) atom_add(&c->live_banks, -1);
    }
}

// ---- task queues ------------------------------------------------------------

// A range of parents of one query at one depth. The first parent starts at
// candidate `skip` of its driver row and the last one stops at `last_end`
// (ALL_CANDIDATES: its whole row), which lets a task hold part of a row.
constexpr uint32_t ALL_CANDIDATES = 0xFFFFFFFFu;
// TaskDesc::q = query slot | tag << 24. Serving tags a query's tasks with the
// low byte of its ticket, so a task that outlives its query is recognised
// (the slot is reused by later queries); batch launches use tag 0.
// Idle waves back off exponentially between polls of the queues, up to
// 2^FM_IDLE_BACKOFF_LOG2 sleeps of 8 x 64 cycles.
#ifndef FM_IDLE_BACKOFF_LOG2
#define FM_IDLE_BACKOFF_LOG2 5
#endif
constexpr uint32_t QSLOT_MASK = 0xFFFFFFu;
__device__ __forceinline__ uint32_t qslot(uint32_t q) { return q & QSLOT_MASK; }
struct TaskDesc {
    uint32_t q, depth, bank, begin, count, skip, last_end;
};

// Pushes `m` tasks (lane i < m describes task i) at priority level lvl.
// All lanes call. Row data the tasks reference must be written before.
__device__ __forceinline__ void push_tasks(const KParams &P, uint32_t lvl, uint32_t m, const TaskDesc &t, uint32_t bank,
                           unsigned long long *stats) {
    if (m == 0) return;
    const uint32_t lane = lane_id();
    Ctl *c = P.ctl;
    uint32_t pos0 = 0, full = 0;
    if (lane == 0) {
        atom_add(&c->pending, (long long)m);
        if (P.serving) atom_add(&P.qs[qslot(t.q)].qpending, (long long)m);
        if (bank != ROOT_BANK) atom_add(&P.bank_ref[bank], (int32_t)m);
        pos0 = atom_add(&c->lq[lvl].tail, m);
        uint32_t head = ld_rlx(&c->lq[lvl].head);
        // signed: consumers may already have claimed positions past ours
        if ((int32_t)(pos0 + m - head) > (int32_t)P.qcap) full = 1;
        stats[ST_TASKS] += m;
    }
    pos0 = bcast(pos0, 0);
    full = bcast(full, 0);
    if (full) {
        if (lane == 0) raise_error(P, FM_ERR_QUEUE_FULL);
        return;
    }
    // make the rows written by every lane visible before publication
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (lane < m) {
        uint32_t pos = pos0 + lane;
        Slot *s = &P.ring[(size_t)lvl * P.qcap + (pos & (P.qcap - 1))];
        if (wait_eq(&s->seq, pos, c, 13)) {
        s->q = t.q;
        s->depth = t.depth;
        s->bank = t.bank;
        s->begin = t.begin;
        s->count = t.count;
        s->skip = t.skip;
        s->last_end = t.last_end;
        st_rel(&s->seq, pos + 1);
        }
    }
    if (lane == 0) {
        atom_add(&c->lq[lvl].avail, (int32_t)m);
        atom_add(&c->queued, (int32_t)m);
    }
}

// Lane 0 only: takes the task of the lowest level (fewest remaining depths).
__device__ __forceinline__ bool pop_task(const KParams &P, TaskDesc &t) {
    Ctl *c = P.ctl;
    for (uint32_t lvl = 0; lvl < P.nlevels; lvl++) {
        if (ld_rlx(&c->lq[lvl].avail) <= 0) continue;
        if (atom_add(&c->lq[lvl].avail, -1) <= 0) {
            atom_add(&c->lq[lvl].avail, 1);
            continue;
        }
        atom_add(&c->queued, -1);
        uint32_t pos = atom_add(&c->lq[lvl].head, 1u);
        Slot *s = &P.ring[(size_t)lvl * P.qcap + (pos & (P.qcap - 1))];
        if (!wait_eq(&s->seq, pos + 1, c, 14)) return false;
        t.q = s->q;
        t.depth = s->depth;
        t.bank = s->bank;
        t.begin = s->begin;
        t.count = s->count;
        t.skip = s->skip;
        t.last_end = s->last_end;
        st_rel(&s->seq, pos + P.qcap);
        return true;
    }
    return false;
}

// Lane 0 only: claims the next chunk of root vertices.
// A wave's first claim is chunk (workgroup id), without an atomic: the
// host starts root_cursor at the grid size, so later claims take the chunks
// past the first round (thousands of waves starting together would otherwise
// queue on the cursor).
__device__ __forceinline__ bool claim_root(const KParams &P, TaskDesc &t, unsigned long long *stats, bool first_claim) {
    Ctl *c = P.ctl;
    unsigned long long k = blockIdx.x;
    if (!first_claim) {
        if (ld_rlx(&c->root_cursor) >= P.total_root_chunks) return false;
        k = atom_add(&c->root_cursor, 1ull);
    }
    if (k >= P.total_root_chunks) return false;
    // the query whose chunk range holds k
    uint32_t lo = 0, hi = P.nq - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (P.qs[mid].root_chunk_begin <= k) lo = mid;
        else hi = mid - 1;
    }
    const DPlan &pl = P.plans[lo];
    const uint32_t per = P.qs[lo].chunk_roots;
    uint32_t local = (uint32_t)k - P.qs[lo].root_chunk_begin;
    uint32_t first = local * per;
    t.q = lo;
    t.depth = 0;
    t.bank = ROOT_BANK;
    t.begin = first;
    t.count = min(per, pl.nroots - first);
    t.skip = 0;
    t.last_end = ALL_CANDIDATES;
    stats[ST_ROOTS]++;
    return true;
}

// ---- output pages ------------------------------------------------------------
// Each wave owns one output page of its current query and appends rows
// without atomics; a page it stops using while partly filled is handed to
// the next wave of the same query through DQuery::spare. The table of a
// query is the union of its pages (page_owner / page_fill), row order
// unspecified.

struct WaveTab {
    uint32_t q, page, fill, fill0, cap;
    // 1 + the query whose table this wave found full (no page for it): in a
    // batch launch pages are never returned to the arena, so later rows of
    // that query are dropped without touching its page counters. After its
    // table fills, a query with a row cap keeps counting until its count
    // cap stops it (on dense R-MAT 2^24 patterns, 10^10 embeddings past a
    // 2.5x10^7-row cap); each emit retrying a page claim and flagging the
    // table there took the 64-query stress launch from ~1 s to 75 s.
    uint32_t full;
};

// Lane 0 only.
__device__ __forceinline__ void tab_release(const KParams &P, WaveTab &w) {
    if (w.page == NO_PAGE) return;
    DQuery &Q = P.qs[w.q];
    st_rlx(&P.page_fill[w.page], w.fill);
    if (w.fill > w.fill0) atom_add(&Q.rows, (unsigned long long)(w.fill - w.fill0));
    if (w.fill < w.cap) {
        unsigned long long v = ((unsigned long long)(w.page + 1) << 32) | w.fill, zero = 0;
        __hip_atomic_compare_exchange_strong(&Q.spare, &zero, v, __ATOMIC_RELEASE, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
    }
    w.page = NO_PAGE;
}

// Lane 0 only: publishes the rows appended since the last sync (page fill
// and the query's row count) and keeps the page.
__device__ __forceinline__ void tab_sync(const KParams &P, WaveTab &w) {
    if (w.page == NO_PAGE || w.fill == w.fill0) return;
    st_rlx(&P.page_fill[w.page], w.fill);
    atom_add(&P.qs[w.q].rows, (unsigned long long)(w.fill - w.fill0));
    w.fill0 = w.fill;
}

// Lane 0, at the end of a final-level task. A batch launch keeps the page
// for the wave's next task of the same query, so a query uses about
// rows / page_rows + waves pages however small its tasks are. Serving
// tables are dense (emit_rows_dense) and hold no page in the wave.
__device__ __forceinline__ void tab_task_end(const KParams &P, WaveTab &w) {
    if (!P.serving) tab_sync(P, w);
}

// Lane 0 only. Takes the query's spare page (if allowed) or claims a fresh
// one; false when the query's page budget or the arena is exhausted.
__device__ __forceinline__ bool tab_acquire(const KParams &P, WaveTab &w, uint32_t q, uint32_t cap, bool allow_spare,
                            unsigned long long *stats) {
    DQuery &Q = P.qs[q];
    w.q = q;
    w.cap = cap;
    w.page = NO_PAGE;
    if (allow_spare && ld_rlx(&Q.spare) != 0) {
        unsigned long long v = __hip_atomic_exchange(&Q.spare, 0ull, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        if (v) {
            w.page = (uint32_t)(v >> 32) - 1;
            w.fill = w.fill0 = (uint32_t)v;
            return true;
        }
    }
    if (ld_rlx(&P.ctl->page_bump) >= P.npages) return false;
    if (atom_add(&Q.npages, 1u) >= Q.max_pages) return false;
    const uint32_t id = atom_add(&P.ctl->page_bump, 1u);
    if (id >= P.npages) return false;
    st_rlx(&P.page_owner[id], q);
    w.page = id;
    w.fill = w.fill0 = 0;
    stats[ST_PAGES]++;
    return true;
}

// Lane 0, serving: a page of the arena, a recycled one first; NO_PAGE when
// the arena is exhausted.
__device__ __forceinline__ uint32_t srv_page_alloc(const KParams &P, unsigned long long *stats) {
    Ctl *c = P.ctl;
    uint32_t id = NO_PAGE;
    if (ld_rlx(&c->pf_avail) > 0) {
        if (atom_add(&c->pf_avail, -1) > 0) {
            const uint32_t pos = atom_add(&c->pf_head, 1u);
            FSlot *fs = &P.pfree_ring[pos & (P.pfcap - 1)];
            if (!wait_eq(&fs->seq, pos + 1, c, 15)) return NO_PAGE;
            id = fs->id;
            st_rel(&fs->seq, pos + P.pfcap);
        } else {
            atom_add(&c->pf_avail, 1);
        }
    }
    if (id == NO_PAGE && ld_rlx(&c->srv_page_bump) < P.npages) {
        id = atom_add(&c->srv_page_bump, 1u);
        if (id >= P.npages) id = NO_PAGE;
    }
    if (id != NO_PAGE) stats[ST_PAGES]++;
    return id;
}

// Lane 0, serving: the page of slot entry k once its owner has resolved it
// (a page id or FAILED_PAGE).
__device__ __forceinline__ uint32_t wait_page(const KParams &P, uint32_t q, uint32_t k) {
    const uint32_t *e = &P.slot_pages[(size_t)q * P.max_slot_pages + k];
    uint32_t it = 0, v;
    while ((v = __hip_atomic_load(e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) == NO_PAGE)
        if (!spin_ok(P.ctl, it, 18, v, k)) return FAILED_PAGE;
    return v;
}

// All lanes, serving. A query's table is dense: a wave claims cnt rows at
// the query's cursor, so row r lands in slot page r / page_rows at
// r % page_rows, and the query holds ceil(rows / page_rows) pages however
// many waves emit its rows. The wave whose rows include the first row of a
// page allocates it; waves with rows further into that page wait for the
// slot entry. When the arena runs out the query's table ends at the start
// of the first page that failed (trunc_at): every row below it is stored.
template <class F>
__device__ __forceinline__ void emit_rows_dense(const KParams &P, uint32_t q, uint32_t n, bool ok, uint32_t rank, uint32_t cnt,
                                unsigned long long *stats, F write_row) {
    const uint32_t lane = lane_id();
    DQuery &Q = P.qs[q];
    const uint32_t pr = Q.page_rows;
    uint32_t off = 0, keep = 0, k0 = 0, pA = FAILED_PAGE, pB = FAILED_PAGE;
    if (lane == 0) {
        const unsigned long long r0 = atom_add(&Q.cursor, (unsigned long long)cnt);
        // the page whose first row is ours, if any: resolved whatever we keep,
        // since waves that read an older trunc_at may wait for it. Pages
        // resolve in index order (the owner of page k waits for page k - 1,
        // whose owner claimed its rows earlier; the chain ends at page 0), so
        // an exhausted arena truncates at the first page it cannot hold and
        // no page past it takes one of the last free pages.
        const unsigned long long kf = (r0 + pr - 1) / pr;
        if (kf * pr < r0 + cnt && kf < Q.max_pages) {
            const uint32_t prev = kf ? wait_page(P, q, (uint32_t)kf - 1) : 0u;
            uint32_t id = FAILED_PAGE;
            if (prev != FAILED_PAGE && kf * pr < ld_rlx(&Q.trunc_at)) {
                id = srv_page_alloc(P, stats);
                if (id == NO_PAGE) {
                    id = FAILED_PAGE;
                    __hip_atomic_fetch_min(&Q.trunc_at, kf * pr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
            __hip_atomic_fetch_max(&Q.npages, (uint32_t)kf + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            st_rel(&P.slot_pages[(size_t)q * P.max_slot_pages + kf], id);
        }
        // trunc_at only falls: rows dropped here lie at or above its final value
        const unsigned long long lim = ld_rlx(&Q.trunc_at);
        keep = r0 >= lim ? 0u : (uint32_t)min((unsigned long long)cnt, lim - r0);
        if (keep < cnt)
            __hip_atomic_fetch_or(&Q.flags, FM_RESULT_TRUNCATED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (keep) {
            k0 = (uint32_t)(r0 / pr);
            off = (uint32_t)(r0 - (unsigned long long)k0 * pr);
            const uint32_t k1 = (uint32_t)((r0 + keep - 1) / pr);
            pA = wait_page(P, q, k0);
            pB = k1 != k0 ? wait_page(P, q, k1) : pA;
        }
    }
    keep = bcast(keep, 0);
    if (keep == 0) return;
    off = bcast(off, 0);
    pA = bcast(pA, 0);
    pB = bcast(pB, 0);
    if (ok && rank < keep) {
        // rows fit a page or straddle two
        const uint32_t at = off + rank;
        const uint32_t page = at < pr ? pA : pB;
        const uint32_t row = at < pr ? at : at - pr;
        if (page < P.npages) write_row(P.pages + (size_t)page * P.page_words + (size_t)row * n);
    }
}

// All lanes. Appends the rows of the lanes with ok set (rank = their order)
// to query q's table; write_row(dst) stores one row of n words at dst.
template <class F>
__device__ __forceinline__ void emit_rows(const KParams &P, WaveTab &wt, uint32_t q, uint32_t n, bool ok, uint32_t rank, uint32_t cnt,
                          unsigned long long *stats, F write_row) {
    if (P.serving) {
        emit_rows_dense(P, q, n, ok, rank, cnt, stats, write_row);
        return;
    }
    const uint32_t lane = lane_id();
    uint32_t pA = NO_PAGE, fA = 0, nA = 0, pB = NO_PAGE, fB = 0;
    if (lane == 0) {
        const uint32_t cap = (uint32_t)(P.page_words / n);
        if (wt.page != NO_PAGE && wt.q != q) tab_release(P, wt);
        const bool known_full = wt.full == q + 1;
        if (wt.page == NO_PAGE && !known_full) tab_acquire(P, wt, q, cap, true, stats);
        if (wt.page != NO_PAGE) {
            pA = wt.page;
            fA = wt.fill;
            nA = min(cnt, wt.cap - wt.fill);
            wt.fill += nA;
            if (nA < cnt) {
                tab_release(P, wt);
                if (tab_acquire(P, wt, q, cap, false, stats)) {
                    pB = wt.page;
                    fB = wt.fill;
                    wt.fill += cnt - nA;
                }
            }
        }
        if (!known_full && (pA == NO_PAGE || (nA < cnt && pB == NO_PAGE))) {
            __hip_atomic_fetch_or(&P.qs[q].flags, FM_RESULT_TRUNCATED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            wt.full = q + 1;
        }
    }
    pA = bcast(pA, 0);
    fA = bcast(fA, 0);
    nA = bcast(nA, 0);
    pB = bcast(pB, 0);
    fB = bcast(fB, 0);
    if (ok) {
        uint32_t *dst = nullptr;
        if (rank < nA) dst = P.pages + (size_t)pA * P.page_words + (size_t)(fA + rank) * n;
        else if (pB != NO_PAGE) dst = P.pages + (size_t)pB * P.page_words + (size_t)(fB + rank - nA) * n;
        if (dst) write_row(dst);
    }
}

// ---- the matcher --------------------------------------------------------------
// Copies n words from global memory into LDS, eight loads per lane in
// flight per round (a plain copy loop waits for each load before its store).
__device__ __forceinline__ void copy_to_lds(uint32_t *dst, const uint32_t *src, uint32_t n) {
    constexpr uint32_t U = 8;
    const uint32_t lane = lane_id();
    for (uint32_t w0 = 0; w0 < n; w0 += U * WAVE) {
        uint32_t v[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) v[u] = src[min(w0 + u * WAVE + lane, n - 1)];
#pragma unroll
        for (uint32_t u = 0; u < U; u++)
            if (w0 + u * WAVE + lane < n) dst[w0 + u * WAVE + lane] = v[u];
    }
}


struct WaveOut {
    uint32_t bank;       // current output bank or NO_BANK
    uint32_t used;       // words used in the bank
    uint32_t run_begin;  // word offset of the rows not yet pushed
};

// LDS words holding a batch's parent rows and requirement rows (start,
// length); a batch takes min(tile_parents, POOL_WORDS / (W + 2k)) parents of
// width W with k requirement rows each (tile_parents <= 64), so the pool fits
// any plan (W <= 32, k <= 16).
constexpr uint32_t POOL_WORDS = 2048;
static_assert(POOL_WORDS / (MAXN + 2 * MAXR) >= 1, "pool too small for the widest batch");

struct Lds {
    uint32_t pool[POOL_WORDS];          // parents | rows start | rows length
    uint32_t dstart[WAVE];              // driver row start
    alignas(16) uint32_t cum[WAVE];     // inclusive scan of driver lengths
    uint32_t didx[WAVE];                // driver requirement index
    // instrumented launches: per-task sums, flushed once per task (lane 0)
    unsigned long long inst[4 + MAXN];  // row words, lookups, candidates, probes, partials
    // instrumented launches, per parent of the batch: its requirement row
    // words, and flags (bit 0: alive and its driver range starts in this
    // task at candidate 0; bit 1: its driver row is empty). A parent is
    // counted once, by the task whose processed range holds its candidate 0.
    unsigned long long acct_words[WAVE];
    uint32_t acct_flags[WAVE];
};

// Pushes the rows written since run_begin as tasks of at most `per`
// parents each (depth D, row width D + 1).
__device__ __forceinline__ void flush_run(const KParams &P, WaveOut &o, uint32_t q, uint32_t D, uint32_t lvl, uint32_t per,
                          unsigned long long *stats) {
    const uint32_t W = D + 1;
    uint32_t rows = (o.used - o.run_begin) / W;
    if (rows == 0) return;
    if (lane_id() == 0) {
        DQuery &Q = P.qs[qslot(q)];
        unsigned long long stored = atom_add(&Q.stored, (unsigned long long)rows) + rows;
        if (Q.max_partials && stored > Q.max_partials) query_fail(P, qslot(q), FM_ERR_COUNT_LIMIT);
    }
    uint32_t ntasks = (rows + per - 1) / per;
    const uint32_t lane = lane_id();
    for (uint32_t t0 = 0; t0 < ntasks; t0 += WAVE) {
        uint32_t m = min(WAVE, ntasks - t0);
        TaskDesc td;
        uint32_t ti = t0 + lane;
        td.q = q;
        td.depth = D;
        td.bank = o.bank;
        td.begin = o.run_begin + ti * per * W;
        td.count = min(per, rows - min(rows, ti * per));
        td.skip = 0;
        td.last_end = ALL_CANDIDATES;
        push_tasks(P, lvl, m, td, o.bank, stats);
    }
    o.run_begin = o.used;
}


// Count-only last level whose vertex has one anchored requirement, no
// self-loop requirement and no order constraint, for a task below the root
// level of a launch that gathers no counters (a leaf of the matching order:
// trees, and cycles closed at an earlier depth). A candidate then needs no
// membership probe, only the check that it is not a vertex of its parent.
// Batches stage up to LEAF_TILE x tile_parents parents (LEAF_TILE per lane),
// so the staging of a batch (parent rows, one row lookup per parent, the
// scan) is paid once per 256 parents instead of 64; every lane keeps
// LEAF_U driver-row loads in flight. Every candidate is still loaded and
// checked one by one. Splits and the first/last parent ranges follow
// process_task.
// Work splitting: a busy wave donates part of its task while idle waves
// outnumber the waiting work (queued tasks and unclaimed root chunks): with
// thousands of waves idle, every busy wave would otherwise cut its work into
// pieces at every check. The queue and root counters, which every push, pop
// and root claim updates, are read only while some wave is idle: polled by
// every busy wave at every check they slow those updates down (one-query
// launches took up to 1.8x longer). The counters are read at the check:
// deciding on values read one check earlier let more waves donate for the
// same idle waves (up to 1.8x the splits) and cost more than the round trip.
// A donated range is staged again by the wave that takes it (parent rows
// into LDS, one row lookup per parent), so candidates are donated only from
// long enough ranges: 2,048 candidates in inner levels, 4,096 in count-only
// leaf batches. On the plain R-MAT 2^20 workload (32 queries) the thresholds
// of 512 took 283 ms of kernel time per step against 263 ms. Serving keeps
// 512: a query served while few others run needs its hub rows spread over
// the idle waves (at 200 requests/s on R-MAT 2^22, p50 latency 10.7 ms
// against 7.6 ms with the larger threshold).
constexpr uint32_t SPLIT_MIN_CANDIDATES = 32 * WAVE;
// While waves are idle, a busy wave queues its children in tasks of this
// many parents instead of task_rows, so they spread sooner. Every task costs
// atomics on words all waves share: on the plain R-MAT 2^20 workload (32
// queries) tasks of 64 parents took 243 ms of kernel time per step, 256
// took 225 ms, 1,024 took 226 ms. Serving keeps 64, for the latency of a
// query served while few others run.
constexpr uint32_t IDLE_TASK_ROWS = 4 * WAVE;
constexpr uint32_t SPLIT_MIN_CANDIDATES_SERVING = 8 * WAVE;
constexpr uint32_t SPLIT_MIN_CANDIDATES_LEAF = 64 * WAVE;
__device__ __forceinline__ bool split_wanted(const KParams &P) {
    uint32_t want = 0;
    const int32_t idle = lane_id() == 0 ? ld_rlx(&P.ctl->idle_waves) : 0;
    if (lane_id() == 0 && idle > 0) {
        const unsigned long long rc = ld_rlx(&P.ctl->root_cursor);
        const long long roots_left = rc < P.total_root_chunks ? (long long)(P.total_root_chunks - rc) : 0;
        want = (long long)idle > (long long)max(0, ld_rlx(&P.ctl->queued)) + roots_left ? 1u : 0u;
    }
    return bcast(want, 0) != 0;
}

// Lane 0 adds a wave's count to its query's every flush_at embeddings (and
// at task end); the count cap is checked there. With a cap, the wave
// flushes after about 1/4096 of it: every wave counting 2^20 before its
// check let 3,072 waves count ~35x past a 3.9x10^7 cap on dense R-MAT 2^24
// patterns before the query stopped.
__device__ __forceinline__ unsigned long long count_flush_at(const DQuery &Q) {
    const unsigned long long cap = Q.max_count;
    return cap ? min(1ull << 20, max(cap >> 12, 4096ull)) : (1ull << 20);
}

constexpr uint32_t LEAF_TILE = 4;
constexpr uint32_t LEAF_MAX = LEAF_TILE * WAVE;  // parents per batch
constexpr uint32_t LEAF_U = 8;                   // candidates per lane per pass
// pool: 0 (word 3, the sum before parent 0) | cum[LEAF_MAX] | dstart[LEAF_MAX] | parent rows
constexpr uint32_t LEAF_HEAD = 4;
static_assert(LEAF_HEAD + 2 * LEAF_MAX + MAXN <= POOL_WORDS, "leaf batch arrays fit the pool");

__device__ __forceinline__ void process_leaf_task(const KParams &P, const TaskDesc &t, Lds &s, unsigned long long *stats,
                                  unsigned long long deadline) {
    const uint32_t lane = lane_id();
    const uint32_t qi = qslot(t.q);
    const DPlan &pl = P.plans[qi];
    DQuery &Q = P.qs[qi];
    const DGraph &g = P.g;
    const uint32_t n = pl.n, d = t.depth, W = d + 1, D = d + 1;
    const uint32_t req = pl.req[pl.req_begin[D]];
    const uint32_t lvl_self = n - 2 - d;
    uint32_t *const cum = s.pool + LEAF_HEAD;  // 16-byte aligned; cum[-1] == 0
    uint32_t *const dstart = cum + LEAF_MAX;
    uint32_t *const par = dstart + LEAF_MAX;
    const uint32_t NB = min(LEAF_TILE * P.tile_parents, (POOL_WORDS - LEAF_HEAD - 2 * LEAF_MAX) / W);
    const uint32_t naut = pl.naut > 1 ? pl.naut : 1u;
    unsigned long long cnt_acc = 0;
    const unsigned long long flush_at = count_flush_at(Q);
    uint32_t cnt_local = t.count;    // shrinks when the upper parents are donated
    uint32_t last_end = t.last_end;  // end of the last parent's driver range
    for (uint32_t b0 = 0; b0 < cnt_local; b0 += NB) {
        uint32_t stop = 0;
        if (lane == 0) stop = (ld_rlx(&P.ctl->abort) || ld_rlx(&Q.status)) ? 1u : 0u;
        if (bcast(stop, 0)) break;
        if (wall_clock64() > deadline) {
            if (lane == 0) timeout_at(P.ctl, 2, t.q, t.depth, t.count);
            break;
        }
        const uint32_t nb = min(NB, cnt_local - b0);
        copy_to_lds(par, P.banks + (size_t)t.bank * P.bank_words + t.begin + (size_t)b0 * W, nb * W);
        __syncthreads();
        // driver range of each parent (LEAF_TILE per lane) and their scan;
        // entries past the batch hold its total
        uint32_t total = 0;
#pragma unroll
        for (uint32_t r = 0; r < LEAF_TILE; r++) {
            const uint32_t p = r * WAVE + lane;
            uint32_t dl = 0;
            if (p < nb) {
                uint32_t st, len;
                row_of(g, req, par[p * W + req_anchor(req)], st, len);
                uint32_t endv = len;
                if (b0 + p == cnt_local - 1) endv = min(endv, last_end);
                const uint32_t startv = (b0 == 0 && p == 0) ? min(t.skip, endv) : 0u;
                dstart[p] = st + startv;
                dl = endv - startv;
            }
            const uint32_t incl = wave_incl_scan(dl);
            cum[p] = total + incl;
            total += bcast(incl, WAVE - 1);
        }
        if (lane == 0) cum[-1] = 0;
        uint32_t T = total;
        __syncthreads();
        // ends of the batch's four 64-parent groups (the same for every lane)
        uint32_t gend[LEAF_TILE];
#pragma unroll
        for (uint32_t j = 0; j < LEAF_TILE; j++) gend[j] = cum[64 * j + 63];
        uint32_t hits = 0;
        for (uint32_t base = 0, it = 0; base < T; base += LEAF_U * WAVE, it++) {
            // work splitting while waves are idle (as in process_task), every
            // 2,048 candidates
            if ((it & 3u) == 0) {
                if (split_wanted(P)) {
                    const uint32_t rem_par = cnt_local > b0 + nb ? cnt_local - (b0 + nb) : 0;
                    if (rem_par >= 2 * WAVE) {
                        // the upper half of the parents after this batch
                        const uint32_t keep = rem_par / 2, first = b0 + nb + keep;
                        TaskDesc r{t.q, d, t.bank, t.begin + first * W, rem_par - keep, 0, last_end};
                        push_tasks(P, lvl_self, 1, r, t.bank, stats);
                        cnt_local = first;
                        last_end = ALL_CANDIDATES;
                        if (lane == 0) stats[ST_SPLIT]++;
                    } else if (T - base >= SPLIT_MIN_CANDIDATES_LEAF) {
                        // the upper half of this batch's remaining candidates
                        // and the parents after them
                        const uint32_t mid = base + ((T - base) / (2 * WAVE)) * WAVE;
                        uint32_t lo = 0, hi = nb - 1;
                        while (lo < hi) {
                            const uint32_t m2 = (lo + hi) >> 1;
                            if (cum[m2] > mid) hi = m2;
                            else lo = m2 + 1;
                        }
                        const uint32_t excl = lo ? cum[lo - 1] : 0u;
                        const uint32_t abs_mid = (mid - excl) + ((b0 == 0 && lo == 0) ? t.skip : 0u);
                        TaskDesc r{t.q, d, t.bank, t.begin + (b0 + lo) * W, cnt_local - b0 - lo, abs_mid, last_end};
                        push_tasks(P, lvl_self, 1, r, t.bank, stats);
                        cnt_local = b0 + lo + 1;
                        last_end = abs_mid;
                        T = mid;
                        if (lane == 0) stats[ST_SPLIT]++;
                    }
                }
            }
            // The lane's LEAF_U candidates (base + u * 64 + lane) advance in
            // lockstep: each round issues its LDS reads for all of them
            // before one wait, with no branch between rounds. The parent of
            // candidate tt is the number of entries with cum <= tt: 64-parent
            // group (registers), 16-parent group, 4-parent group, parent.
            uint32_t tt[LEAF_U], lo[LEAF_U], r4[LEAF_U][4];
#pragma unroll
            for (uint32_t u = 0; u < LEAF_U; u++) {
                tt[u] = min(base + u * WAVE + lane, T - 1);
                uint32_t l = 0;
#pragma unroll
                for (uint32_t j = 0; j < LEAF_TILE; j++) l += gend[j] <= tt[u] ? 64u : 0u;
                lo[u] = l;
            }
#pragma unroll
            for (uint32_t u = 0; u < LEAF_U; u++)
#pragma unroll
                for (uint32_t j = 0; j < 4; j++) r4[u][j] = cum[lo[u] + 16 * j + 15];
#pragma unroll
            for (uint32_t u = 0; u < LEAF_U; u++)
#pragma unroll
                for (uint32_t j = 0; j < 4; j++) lo[u] += r4[u][j] <= tt[u] ? 16u : 0u;
#pragma unroll
            for (uint32_t u = 0; u < LEAF_U; u++)
#pragma unroll
                for (uint32_t j = 0; j < 4; j++) r4[u][j] = cum[lo[u] + 4 * j + 3];
#pragma unroll
            for (uint32_t u = 0; u < LEAF_U; u++)
#pragma unroll
                for (uint32_t j = 0; j < 4; j++) lo[u] += r4[u][j] <= tt[u] ? 4u : 0u;
            uint32_t ex[LEAF_U];
#pragma unroll
            for (uint32_t u = 0; u < LEAF_U; u++) {
                const uint4 c4 = *reinterpret_cast<const uint4 *>(&cum[lo[u]]);
                r4[u][0] = c4.x;
                r4[u][1] = c4.y;
                r4[u][2] = c4.z;
                r4[u][3] = c4.w;
                ex[u] = cum[(int)lo[u] - 1];  // cum[-1] == 0
            }
#pragma unroll
            for (uint32_t u = 0; u < LEAF_U; u++)
#pragma unroll
                for (uint32_t j = 0; j < 4; j++) {
                    const bool le = r4[u][j] <= tt[u];
                    lo[u] += le ? 1u : 0u;
                    ex[u] = le ? r4[u][j] : ex[u];
                }
            uint32_t xs[LEAF_U];
#pragma unroll
            for (uint32_t u = 0; u < LEAF_U; u++) xs[u] = dstart[lo[u]] + (tt[u] - ex[u]);
#pragma unroll
            for (uint32_t u = 0; u < LEAF_U; u++) xs[u] = g.col[xs[u]];
            bool ok[LEAF_U];
#pragma unroll
            for (uint32_t u = 0; u < LEAF_U; u++) {
                ok[u] = base + u * WAVE + lane < T;
                lo[u] *= W;  // the parent's row in LDS
            }
            for (uint32_t a = 0; a < W; a++) {
                uint32_t pv[LEAF_U];
#pragma unroll
                for (uint32_t u = 0; u < LEAF_U; u++) pv[u] = par[lo[u] + a];
#pragma unroll
                for (uint32_t u = 0; u < LEAF_U; u++) ok[u] &= pv[u] != xs[u];
            }
#pragma unroll
            for (uint32_t u = 0; u < LEAF_U; u++) hits += ok[u] ? 1u : 0u;
        }
        cnt_acc += wave_sum_u64(hits) * naut;
        if (lane == 0 && cnt_acc >= flush_at) {
            const unsigned long long now_cnt = atom_add(&Q.count, cnt_acc) + cnt_acc;
            cnt_acc = 0;
            if (Q.max_count && now_cnt > Q.max_count) query_fail(P, qi, FM_ERR_COUNT_LIMIT);
        }
        __syncthreads();
    }
    if (lane == 0 && cnt_acc) {
        const unsigned long long now_cnt = atom_add(&Q.count, cnt_acc) + cnt_acc;
        if (Q.max_count && now_cnt > Q.max_count) query_fail(P, qi, FM_ERR_COUNT_LIMIT);
    }
}

// Extends one task by one depth. Returns after pushing children tasks; the
// caller completes the task (bank reference and pending count). PAIRED takes
// two candidates per lane per step (more loads in flight, more registers).
template <bool PAIRED>
__device__ __forceinline__ void process_task(const KParams &P, const TaskDesc &t, Lds &s, WaveOut &o, WaveTab &wt,
                             unsigned long long *stats, unsigned long long &starve_since,
                             unsigned long long deadline) {
    const uint32_t lane = lane_id();
    const uint32_t qi = qslot(t.q);
    const DPlan &pl = P.plans[qi];
    DQuery &Q = P.qs[qi];
    if (ld_rlx(&Q.status) != 0) return;
    const bool table = P.table && !Q.count_only;  // this query stores rows
    const uint32_t n = pl.n;
    const uint32_t d = t.depth;
    const uint32_t W = d + 1;  // parent width
    const bool root = t.bank == ROOT_BANK;
    const DGraph &g = P.g;
    // A malformed task (an engine bug) must end the launch with a status,
    // never read out of bounds.
    {
        bool bad = t.count == 0 || d + 1 >= max(n, 2u) || (t.count == 1 && t.skip > t.last_end);
        if (root) bad = bad || d != 0 || (unsigned long long)t.begin + t.count > pl.nroots;
        else bad = bad || t.bank >= P.nbanks || (unsigned long long)t.begin + (unsigned long long)t.count * W > P.bank_words;
        if (bad) {
            if (lane == 0) {
                uint32_t zero = 0;
                if (__hip_atomic_compare_exchange_strong(&P.ctl->error, &zero, (uint32_t)FM_ERR_INTERNAL,
                                                         __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT)) {
                    uint32_t *dg = P.ctl->diag;
                    dg[0] = t.q;
                    dg[1] = t.depth;
                    dg[2] = t.bank;
                    dg[3] = t.begin;
                    dg[4] = t.count;
                    dg[5] = t.skip;
                    dg[6] = t.last_end;
                    dg[7] = n;
                    dg[8] = pl.nroots;
                    dg[9] = ld_rlx(&Q.pad);
                    dg[10] = (uint32_t)__hip_atomic_load(&Q.qpending, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                st_rlx(&P.ctl->abort, 1u);
            }
            return;
        }
    }

    // ---- single-vertex query: every root is an embedding -----------------
    if (n == 1) {
        unsigned long long cnt_acc = 0;
        for (uint32_t b0 = 0; b0 < t.count; b0 += WAVE) {
            uint32_t nb = min(WAVE, t.count - b0);
            bool ok = lane < nb;
            uint32_t v = t.begin + b0 + lane;
            if (ok && pl.root_base != ALL_VERTICES) v = g.roots[pl.root_base + t.begin + b0 + lane];
            for (uint32_t i = 0; i < pl.nself[0] && ok; i++) {
                uint32_t st, len;
                row_of(g, pl.req[pl.req_begin[0] + i], v, st, len);
                ok = row_contains(g.col, g.skip, st, len, v);
            }
            unsigned long long mask = __ballot(ok);
            uint32_t cnt = __popcll(mask);
            uint32_t rank = __popcll(mask & ((1ull << lane) - 1ull));
            cnt_acc += cnt;
            if (P.instrument && lane == 0) s.inst[4 + 0] += cnt;
            if (table && cnt)
                emit_rows(P, wt, qi, 1, ok, rank, cnt, stats, [&](uint32_t *dst) { dst[0] = v; });
        }
        if (lane == 0) {
            atom_add(&Q.count, cnt_acc);
            if (table) tab_task_end(P, wt);
        }
        return;
    }

    const uint32_t D = d + 1;  // depth of the children
    const bool final_level = D == n - 1;
    const uint32_t lvl_children = final_level ? 0 : n - 2 - D;
    const uint32_t lvl_self = n - 2 - d;
    const uint32_t rb = pl.req_begin[D];
    const uint32_t nself = pl.nself[D];
    const uint32_t k = pl.req_begin[D + 1] - rb - nself;  // anchored requirements
    if (final_level && !table && !P.instrument && !root && k == 1 && nself == 0 &&
        pl.sc_begin[D] == pl.sc_begin[D + 1]) {
        process_leaf_task(P, t, s, stats, deadline);
        return;
    }

    // batch layout in the LDS pool
    const uint32_t NB = min(P.tile_parents, POOL_WORDS / (W + 2 * k));
    uint32_t *const par = s.pool;
    uint32_t *const rows_st = s.pool + NB * W;
    uint32_t *const rows_len = rows_st + NB * k;

    // Children go out as tasks of up to task_rows parents; while waves are
    // idle, of one batch (64) so that they spread sooner.
    uint32_t run_rows = P.task_rows;

    unsigned long long cnt_acc = 0;
    const unsigned long long flush_at = count_flush_at(Q);
    uint32_t cnt_local = t.count;       // shrinks when the upper parents are donated
    uint32_t last_end = t.last_end;     // end of the last parent's driver range
    for (uint32_t b0 = 0; b0 < cnt_local; b0 += NB) {
        // periodic checks
        uint32_t stop = 0;
        if (lane == 0) stop = (ld_rlx(&P.ctl->abort) || ld_rlx(&Q.status)) ? 1u : 0u;
        if (bcast(stop, 0)) break;
        if (wall_clock64() > deadline) {
            if (lane == 0) timeout_at(P.ctl, 2, t.q, t.depth, t.count);
            break;
        }
        const uint32_t nb = min(NB, cnt_local - b0);
        // ---- parents into LDS ----
        if (root) {
            if (lane < nb) {
                uint32_t idx = t.begin + b0 + lane;
                par[lane] = pl.root_base == ALL_VERTICES ? idx : g.roots[pl.root_base + idx];
            }
        } else {
            copy_to_lds(par, P.banks + (size_t)t.bank * P.bank_words + t.begin + (size_t)b0 * W, nb * W);
        }
        __syncthreads();
        // ---- requirement rows per parent; driver = shortest ----
        uint32_t dlen = 0, dst_ = 0, didx = 0;
        unsigned long long words = 0;
        bool alive = lane < nb;
        if (alive && root && pl.nself[0]) {
            uint32_t v = par[lane];
            for (uint32_t i = 0; i < pl.nself[0] && alive; i++) {
                uint32_t st, len;
                row_of(g, pl.req[pl.req_begin[0] + i], v, st, len);
                alive = row_contains(g.col, g.skip, st, len, v);
            }
        }
        uint32_t acct = 0;
        if (lane < nb) {
            uint32_t best = 0xFFFFFFFFu;
            for (uint32_t i = 0; i < k; i++) {
                uint32_t r = pl.req[rb + i];
                uint32_t st, len;
                row_of(g, r, par[lane * W + req_anchor(r)], st, len);
                rows_st[lane * k + i] = st;
                rows_len[lane * k + i] = len;
                words += len;
                if (len < best) {
                    best = len;
                    dst_ = st;
                    didx = i;
                }
            }
            uint32_t endv = best;
            if (b0 + lane == cnt_local - 1) endv = min(endv, last_end);
            const bool first_par = b0 == 0 && lane == 0;
            const uint32_t startv = first_par ? min(t.skip, endv) : 0;
            dst_ += startv;
            dlen = alive ? endv - startv : 0;
            acct = (alive && !(first_par && t.skip)) ? 1u : 0u;
            if (best == 0) acct |= 2u;
        }
        s.dstart[lane] = dst_;
        s.didx[lane] = didx;
        const uint32_t incl = wave_incl_scan(dlen);
        s.cum[lane] = incl;
        uint32_t T = bcast(incl, WAVE - 1);
        if (P.instrument) {
            s.acct_words[lane] = words;
            s.acct_flags[lane] = acct;
        }
        __syncthreads();
        // Instrumented launches: the batch's algorithmic counters once its
        // processed candidates are final ([0, Tfin) of the flattened range:
        // a split or a yield hands the rest to another task). A parent's row
        // words, lookups and root record are counted by the one task whose
        // processed range holds its candidate 0 (with an empty driver row:
        // the task that holds the parent), so the sums do not depend on how
        // the work was split.
        auto account = [&](uint32_t Tfin) {
            if (!P.instrument) return;
            const uint32_t fl = s.acct_flags[lane];
            const uint32_t hi = s.cum[lane], lo = lane ? s.cum[lane - 1] : 0u;
            const bool c = (fl & 1u) && ((fl & 2u) ? hi <= Tfin : (lo < hi && lo < Tfin));
            const unsigned long long wsum = wave_sum_u64(c ? s.acct_words[lane] : 0ull);
            const uint32_t nc = __popcll(__ballot(c));
            if (lane == 0) {
                s.inst[0] += wsum;
                s.inst[1] += (unsigned long long)nc * k;
                if (root) s.inst[4 + 0] += nc;
                s.inst[2] += Tfin;
            }
        };
        // ---- candidates, 64 per step ----
        // Each step's candidate word (and its parent) is loaded during the
        // step before, so the load overlaps that step's probes. The position
        // is clamped to the batch's last candidate, so the load is always in
        // bounds and unconditional (no loop-carried value merged at a
        // branch); T only shrinks, so a prefetched position below the
        // current T is that step's candidate.
        // The parent of candidate tt is the number of parents p with
        // cum[p] <= tt: cum is nondecreasing and every one of its 64 entries
        // is written per batch, those past the batch's parents holding its
        // original total, which exceeds tt. Three rounds of 4 independent
        // LDS reads (16-parent groups, 4-parent groups, parents) find it,
        // instead of a chain of 6 reads in a binary search.
        auto locate = [&](uint32_t from, uint32_t &pi_out, uint32_t &x_out) {
            const uint32_t tt = min(from + lane, T - 1);
            uint32_t lo = 0;
#pragma unroll
            for (uint32_t j = 0; j < 4; j++) lo += s.cum[16 * j + 15] <= tt ? 16u : 0u;
            uint32_t lo4 = lo;
#pragma unroll
            for (uint32_t j = 0; j < 4; j++) lo4 += s.cum[lo + 4 * j + 3] <= tt ? 4u : 0u;
            const uint4 c4 = *reinterpret_cast<const uint4 *>(&s.cum[lo4]);
            uint32_t excl = lo4 ? s.cum[lo4 - 1] : 0u;
            lo = lo4;
            if (c4.x <= tt) { lo++; excl = c4.x; }
            if (c4.y <= tt) { lo++; excl = c4.y; }
            if (c4.z <= tt) { lo++; excl = c4.z; }
            if (c4.w <= tt) { lo++; excl = c4.w; }
            pi_out = lo;
            x_out = g.col[s.dstart[lo] + (tt - excl)];
        };
        // Work splitting: when waves are idle, donate half of what is
        // left: the upper half of the remaining parents when at least
        // two batches remain, else the upper half of this batch's
        // remaining candidates together with the parents after them (a
        // hub parent's long driver row is split across waves).
        auto maybe_split = [&](uint32_t base) __attribute__((always_inline)) {
            const uint32_t want = split_wanted(P) ? 1u : 0u;
            run_rows = want ? min(P.task_rows, PAIRED ? IDLE_TASK_ROWS : WAVE) : P.task_rows;
            if (want) {
                // a candidate split may already have ended the task inside this batch
                const uint32_t rem_par = cnt_local > b0 + nb ? cnt_local - (b0 + nb) : 0;
                if (rem_par >= 2 * WAVE) {
                    const uint32_t keep = rem_par / 2, first = b0 + nb + keep;
                    TaskDesc r{t.q, d, t.bank, root ? t.begin + first : t.begin + first * W, rem_par - keep, 0,
                               last_end};
                    push_tasks(P, lvl_self, 1, r, t.bank, stats);
                    cnt_local = first;
                    last_end = ALL_CANDIDATES;
                    if (lane == 0) stats[ST_SPLIT]++;
                } else if (T - base >= (PAIRED ? SPLIT_MIN_CANDIDATES : SPLIT_MIN_CANDIDATES_SERVING)) {
                    const uint32_t mid = base + ((T - base) / (2 * WAVE)) * WAVE;
                    uint32_t lo = 0, hi = nb - 1;
                    while (lo < hi) {
                        uint32_t m2 = (lo + hi) >> 1;
                        if (s.cum[m2] > mid) hi = m2;
                        else lo = m2 + 1;
                    }
                    const uint32_t excl = lo ? s.cum[lo - 1] : 0;
                    // driver-row offset of candidate `mid` in parent lo
                    const uint32_t abs_mid = (mid - excl) + ((b0 == 0 && lo == 0) ? t.skip : 0);
                    TaskDesc r{t.q, d, t.bank, root ? t.begin + b0 + lo : t.begin + (b0 + lo) * W,
                               cnt_local - b0 - lo, abs_mid, last_end};
                    push_tasks(P, lvl_self, 1, r, t.bank, stats);
                    // this task now ends inside parent lo
                    cnt_local = b0 + lo + 1;
                    last_end = abs_mid;
                    T = mid;
                    if (lane == 0) stats[ST_SPLIT]++;
                }
            }
        };
        // Count-only last level whose vertex has one requirement (the last
        // vertex of the order is a leaf: trees, and cycles closed at an
        // earlier depth): a candidate needs no membership probe, only the
        // check that it is not a vertex of its parent (LDS). Every lane then
        // keeps LEAF_U driver-row loads in flight per iteration, where the
        // general loop below has one step prefetched. Every candidate is
        // still loaded and checked one by one.
        const bool leaf_count =
            final_level && !table && k == 1 && nself == 0 && pl.sc_begin[D] == pl.sc_begin[D + 1];
        if (leaf_count) {
            constexpr uint32_t LEAF_U = 8;
            uint32_t hits = 0;
            for (uint32_t base = 0, it = 0; base < T; base += LEAF_U * WAVE, it++) {
                if ((it & 1u) == 0) maybe_split(base);  // every 1,024 candidates (T stays > base)
                uint32_t xs[LEAF_U], ps[LEAF_U];
#pragma unroll
                for (uint32_t u = 0; u < LEAF_U; u++) locate(base + u * WAVE, ps[u], xs[u]);
#pragma unroll
                for (uint32_t u = 0; u < LEAF_U; u++) {
                    const uint32_t *const prow = &par[ps[u] * W];
                    bool ok = base + u * WAVE + lane < T;
                    for (uint32_t a = 0; a < W; a++) ok &= prow[a] != xs[u];
                    hits += ok ? 1u : 0u;
                }
            }
            const unsigned long long nh = wave_sum_u64(hits);
            if (P.instrument && lane == 0) s.inst[4 + D] += nh;
            cnt_acc += nh * (pl.naut > 1 ? pl.naut : 1u);
        }
        // Each step takes STEP candidates: one per lane, or with PAIRED two
        // (base + lane and base + 64 + lane) whose probes are in flight
        // together.
        constexpr uint32_t STEP = PAIRED ? 2 * WAVE : WAVE;
        uint32_t pf_pi = 0, pf_x = 0, pf_pi2 = 0, pf_x2 = 0;
        if (T && !leaf_count) {
            locate(0, pf_pi, pf_x);
            if constexpr (PAIRED) locate(WAVE, pf_pi2, pf_x2);
        }
        for (uint32_t base = 0, step = 0; base < T && !leaf_count; base += STEP, step++) {
            if ((step & (PAIRED ? 1u : 3u)) == 0) maybe_split(base);  // every 256 candidates
            if (!final_level) {
                const uint32_t need = STEP * (D + 1);
                static_assert(2 * WAVE * MAXN <= 2 * MIN_BANK_SIZE, "a fresh bank holds one step of children");
                bool yield = false;
                if (o.bank != NO_BANK && o.used + need > P.bank_words) {
                    // The output bank is full: hand its rows over as tasks and
                    // yield, so the deeper tasks just pushed run before this
                    // one produces more (bounded, depth-first memory use).
                    flush_run(P, o, t.q, D, lvl_children, run_rows, stats);
                    if (lane == 0) bank_release(P, o.bank);
                    o.bank = NO_BANK;
                    yield = true;
                } else if (o.bank == NO_BANK) {
                    uint32_t nbk = NO_BANK;
                    if (lane == 0) nbk = bank_alloc(P, stats);
                    nbk = bcast(nbk, 0);
                    if (nbk == NO_BANK) {
                        // No bank: requeue and take deeper (bank-freeing) work.
                        if (lane == 0) {
                            stats[ST_ALLOC_FAIL]++;
                            unsigned long long now = wall_clock64();
                            if (starve_since == 0) starve_since = now;
                            else if (now - starve_since > P.starve_ticks) raise_error(P, FM_ERR_POOL_EXHAUSTED);
                        }
                        yield = true;
                    } else {
                        o.bank = nbk;
                        o.used = 0;
                        o.run_begin = 0;
                        if (lane == 0) starve_since = 0;
                    }
                }
                if (yield) {
                    // remainder: from the parent of candidate `base`
                    uint32_t lo = 0, hi = nb - 1;
                    while (lo < hi) {
                        uint32_t mid = (lo + hi) >> 1;
                        if (s.cum[mid] > base) hi = mid;
                        else lo = mid + 1;
                    }
                    const uint32_t excl = lo ? s.cum[lo - 1] : 0;
                    TaskDesc r;
                    r.q = t.q;
                    r.depth = d;
                    r.bank = t.bank;
                    r.begin = root ? t.begin + b0 + lo : t.begin + (b0 + lo) * W;
                    r.count = cnt_local - b0 - lo;
                    r.skip = (base - excl) + ((b0 == 0 && lo == 0) ? t.skip : 0);
                    r.last_end = last_end;
                    if (lane == 0) stats[ST_REQUEUE]++;
                    account(base);
                    push_tasks(P, lvl_self, 1, r, t.bank, stats);
                    return;
                }
            }
            const uint32_t ta = base + lane, tb = ta + WAVE;
            uint32_t nprobe = 0;
            const uint32_t xa = pf_x, pia = pf_pi, xb = pf_x2, pib = pf_pi2;
            locate(base + STEP, pf_pi, pf_x);
            if constexpr (PAIRED) locate(base + STEP + WAVE, pf_pi2, pf_x2);
            const uint32_t *const prowa = &par[pia * W];
            const uint32_t *const prowb = &par[pib * W];
            // checks without global reads: not a vertex of the parent, and
            // the symmetry-breaking order constraints
            auto local_ok = [&](const uint32_t *prow, uint32_t x) {
                bool ok = true;
                for (uint32_t a = 0; a < W; a++) ok &= prow[a] != x;
                for (uint32_t i = pl.sc_begin[D]; i < pl.sc_begin[D + 1] && ok; i++) {
                    const uint32_t c = pl.sc[i], v = prow[c & 63u];
                    ok = (c & 128u) ? x > v : x < v;
                }
                return ok;
            };
            bool oka = ta < T && local_ok(prowa, xa);
            bool okb = PAIRED && tb < T && local_ok(prowb, xb);
            // anchored requirements other than the driver
            const uint32_t dxa = s.didx[pia];
            if constexpr (PAIRED) {
                const uint32_t dxb = s.didx[pib];
                for (uint32_t i = 0; i < k && (oka || okb); i++) {
                    const bool pa = oka && i != dxa, pb = okb && i != dxb;
                    if (!(pa || pb)) continue;
                    const uint32_t sta = pa ? rows_st[pia * k + i] : 0u, lena = pa ? rows_len[pia * k + i] : 0u;
                    const uint32_t stb = pb ? rows_st[pib * k + i] : 0u, lenb = pb ? rows_len[pib * k + i] : 0u;
                    bool ha, hb;
                    row_contains2(g.col, g.skip, sta, lena, xa, stb, lenb, xb, ha, hb);
                    if (pa) {
                        oka = ha;
                        nprobe++;
                    }
                    if (pb) {
                        okb = hb;
                        nprobe++;
                    }
                }
            } else {
                for (uint32_t i = 0; i < k && oka; i++)
                    if (i != dxa) {
                        oka = row_contains(g.col, g.skip, rows_st[pia * k + i], rows_len[pia * k + i], xa);
                        nprobe++;
                    }
            }
            for (uint32_t i = 0; i < nself && oka; i++) {
                uint32_t st, len;
                row_of(g, pl.req[rb + k + i], xa, st, len);
                oka = row_contains(g.col, g.skip, st, len, xa);
            }
            for (uint32_t i = 0; i < nself && okb; i++) {
                uint32_t st, len;
                row_of(g, pl.req[rb + k + i], xb, st, len);
                okb = row_contains(g.col, g.skip, st, len, xb);
            }
            if (P.instrument) {
                unsigned long long np = wave_sum_u64(nprobe);
                if (lane == 0) s.inst[3] += np;
            }
            const unsigned long long ma = __ballot(oka), mb = PAIRED ? __ballot(okb) : 0ull;
            const uint32_t cna = __popcll(ma), cnb = __popcll(mb);
            if (cna + cnb == 0) continue;
            const unsigned long long below = (1ull << lane) - 1ull;
            const uint32_t ranka = __popcll(ma & below), rankb = __popcll(mb & below);
            if (P.instrument && lane == 0) s.inst[4 + D] += cna + cnb;
            if (final_level) {
                const uint32_t naut = pl.naut > 1 ? pl.naut : 1u;
                cnt_acc += (unsigned long long)(cna + cnb) * naut;
                if (table) {
                    // all lanes call (wave-uniform conditions)
                    auto emit = [&](bool ok, uint32_t rank, uint32_t cnt, const uint32_t *prow, uint32_t x) {
                        if (naut == 1) {
                            emit_rows(P, wt, qi, n, ok, rank, cnt, stats, [&](uint32_t *dst) {
                                for (uint32_t j = 0; j < W; j++) dst[pl.order[j]] = prow[j];
                                dst[pl.order[D]] = x;
                            });
                        } else {
                            // one row per automorphism p: column u holds m(p(u))
                            for (uint32_t ai = 0; ai < naut; ai++)
                                emit_rows(P, wt, qi, n, ok, rank, cnt, stats, [&](uint32_t *dst) {
                                    for (uint32_t u = 0; u < n; u++) {
                                        const uint32_t dd = pl.autcol[ai][u];
                                        dst[u] = dd == D ? x : prow[dd];
                                    }
                                });
                        }
                    };
                    // one call site for both halves keeps emit_rows inlined once
#pragma nounroll
                    for (uint32_t h = 0; h < (PAIRED ? 2u : 1u); h++) {
                        const uint32_t cnt = h ? cnb : cna;
                        if (cnt) emit(h ? okb : oka, h ? rankb : ranka, cnt, h ? prowb : prowa, h ? xb : xa);
                    }
                }
            } else {
                const uint32_t CW = D + 1;
                uint32_t *const out = P.banks + (size_t)o.bank * P.bank_words + o.used;
                if (oka) {
                    uint32_t *dst = out + ranka * CW;
                    for (uint32_t j = 0; j < W; j++) dst[j] = prowa[j];
                    dst[W] = xa;
                }
                if (okb) {
                    uint32_t *dst = out + (cna + rankb) * CW;
                    for (uint32_t j = 0; j < W; j++) dst[j] = prowb[j];
                    dst[W] = xb;
                }
                o.used += (cna + cnb) * CW;
                if ((o.used - o.run_begin) / CW >= run_rows) flush_run(P, o, t.q, D, lvl_children, run_rows, stats);
            }
        }
        account(T);
        if (final_level && lane == 0 && cnt_acc >= flush_at) {
            unsigned long long now_cnt = atom_add(&Q.count, cnt_acc) + cnt_acc;
            cnt_acc = 0;
            if (Q.max_count && now_cnt > Q.max_count) query_fail(P, qi, FM_ERR_COUNT_LIMIT);
        }
        __syncthreads();
    }
    if (!final_level && o.bank != NO_BANK) flush_run(P, o, t.q, D, lvl_children, run_rows, stats);
    // the query's stored-row count is complete when its last task finishes
    if (final_level && table && lane == 0) tab_task_end(P, wt);
    if (lane == 0 && cnt_acc) {
        unsigned long long now_cnt = atom_add(&Q.count, cnt_acc) + cnt_acc;
        if (Q.max_count && now_cnt > Q.max_count) query_fail(P, qi, FM_ERR_COUNT_LIMIT);
    }
}

// ---- serving: admission and completion -------------------------------------

__device__ __forceinline__ uint32_t ld_sys(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void st_sys(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Lane 0: publishes a finished query to the host. Every task of the query
// released its rows before its pending decrement, which this wave acquired.
__device__ __forceinline__ void complete_query(const KParams &P, uint32_t q) {
    DQuery &Q = P.qs[q];
    SrvDone *d = &P.srv_done[q];
    const unsigned long long count = ld_rlx(&Q.count);
    // the table: rows [0, rows) in slot pages 0, 1, ... (dense)
    const unsigned long long rows = min(ld_rlx(&Q.cursor), ld_rlx(&Q.trunc_at));
    const uint32_t pages = Q.page_rows ? (uint32_t)((rows + Q.page_rows - 1) / Q.page_rows) : 0u;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");  // table rows visible to the host
    st_sys(&d->count_lo, (uint32_t)count);
    st_sys(&d->count_hi, (uint32_t)(count >> 32));
    st_sys(&d->rows_lo, (uint32_t)rows);
    st_sys(&d->rows_hi, (uint32_t)(rows >> 32));
    st_sys(&d->status, ld_rlx(&Q.status));
    st_sys(&d->flags, __hip_atomic_load(&Q.flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    st_sys(&d->npages, pages);
    __hip_atomic_store(&d->ticket, Q.pad, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_fetch_add(&P.ctl->srv_state, -SRV_ACTIVE_ONE, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    st_sys(&P.srv->live_banks, (uint32_t)max(0, ld_rlx(&P.ctl->live_banks)));
}

// Lane 0: claims the next submission; returns its slot or NO_PAGE.
__device__ __forceinline__ uint32_t claim_submission(const KParams &P) {
    Ctl *c = P.ctl;
    // The two relaxed loads may complete in either order, so the tail read
    // can be older than the head: the head may already be past it. Only a
    // head strictly behind the tail names a written entry.
    unsigned long long st = ld_rlx(&c->srv_state);
    if (st & SRV_CLOSING) return NO_PAGE;
    const uint32_t h = srv_head(st);
    if ((int32_t)(ld_sys(&P.srv->sub_tail) - h) <= 0) return NO_PAGE;
    // the head advances and the query counts as active in one step; a
    // launch that started closing refuses (its waves are leaving)
    const unsigned long long nst = (st & ~0xFFFFFFFFull) + SRV_ACTIVE_ONE + (h + 1u);
    if (!__hip_atomic_compare_exchange_strong(&c->srv_state, &st, nst, __ATOMIC_ACQ_REL, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT))
        return NO_PAGE;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
    return ld_sys(&P.srv->sub[h % SUB_CAP].slot);
}

// All lanes: admits the query in `slot`: recycles the pages of the slot's
// previous query, copies the plan into device memory, initialises the
// query state and queues its root chunks.
__device__ __forceinline__ void admit(const KParams &P, uint32_t slot, unsigned long long *stats) {
    const uint32_t lane = lane_id();
    DQuery &Q = P.qs[slot];
    Ctl *c = P.ctl;
    // pages of the previous query in this slot go back to the free ring
    // (each entry is taken with an exchange, so no page is freed twice)
    const uint32_t old = min(Q.npages, Q.max_pages);
    for (uint32_t i = lane; i < old; i += WAVE) {
        const uint32_t id = __hip_atomic_exchange(&P.slot_pages[(size_t)slot * P.max_slot_pages + i], NO_PAGE,
                                                  __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (id == NO_PAGE || id >= P.npages) continue;
        uint32_t pos = atom_add(&c->pf_tail, 1u);
        FSlot *fs = &P.pfree_ring[pos & (P.pfcap - 1)];
        if (wait_eq(&fs->seq, pos, c, 17)) {
            fs->id = id;
            st_rel(&fs->seq, pos + 1);
            atom_add(&c->pf_avail, 1);
        }
    }
    // plan: pinned host -> device
    const uint32_t *src = (const uint32_t *)&P.srv_plans[slot];
    uint32_t *dst = (uint32_t *)&P.plans_rw[slot];
    for (uint32_t i = lane; i < sizeof(DPlan) / 4; i += WAVE) dst[i] = ld_sys(src + i);
    const SrvQInit *qi = &P.srv_qinit[slot];
    // query state: zero, then the fields the submission sets (lane 0 stores
    // after the wave's zero stores to the same words)
    static_assert(sizeof(DQuery) % 4 == 0, "DQuery is cleared by words");
    uint32_t *qw = reinterpret_cast<uint32_t *>(&Q);
    for (uint32_t i = lane; i < sizeof(DQuery) / 4; i += WAVE) qw[i] = 0;
    if (lane == 0) {
        const unsigned long long max_rows = ((unsigned long long)ld_sys((const uint32_t *)&qi->max_rows + 1) << 32) |
                                            ld_sys((const uint32_t *)&qi->max_rows);
        Q.max_rows = max_rows;
        Q.trunc_at = max_rows;
        Q.page_rows = ld_sys(&qi->page_rows);
        Q.max_pages = ld_sys(&qi->max_pages);
        Q.pad = ld_sys(&qi->ticket);  // ticket, echoed in the completion
        Q.qpending = 1;               // held until every root chunk is queued
    }
    __syncthreads();
    // publish the plan and the query state; also drop this CU's L1 copies of
    // the slot's old plan before any task of the new query runs here
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
    // The plan's size fields come from the pinned source, not from the copy
    // just stored: the copy's lines may sit stale in this CU's vector L1
    // from the slot's previous query.
    const uint32_t n = ld_sys(&P.srv_plans[slot].n);
    const uint32_t nroots = ld_sys(&P.srv_plans[slot].nroots);
    const uint32_t chunks = (nroots + P.root_chunk - 1) / P.root_chunk;
    const uint32_t lvl = max(n, 2u) - 2;
    for (uint32_t c0 = 0; c0 < chunks; c0 += WAVE) {
        const uint32_t m = min(WAVE, chunks - c0);
        const uint32_t first = (c0 + lane) * P.root_chunk;
        const uint32_t tag = ld_sys(&P.srv_qinit[slot].ticket) & 0xFFu;
        TaskDesc t{slot | tag << 24, 0, ROOT_BANK, first, lane < m ? min(P.root_chunk, nroots - first) : 0, 0,
                   ALL_CANDIDATES};
        push_tasks(P, lvl, m, t, ROOT_BANK, stats);
    }
    if (lane == 0) {
        st_sys(&P.srv->admitted, srv_head(ld_rlx(&c->srv_state)));
        st_sys(&P.srv->live_banks, (uint32_t)max(0, ld_rlx(&c->live_banks)));
        if (atom_add(&Q.qpending, -1ll) == 1) complete_query(P, slot);
    }
}

// The persistent matcher: one wave per workgroup, looping over root chunks,
// queued tasks and (serving) admissions until the launch's work is done.
template <bool PAIRED>
__device__ __forceinline__ void match_waves(const KParams &P) {
    __shared__ Lds s;
    __shared__ TaskDesc s_task;
    __shared__ uint32_t s_flag;
    __shared__ WaveTab wt;  // lane 0 owns it
    const uint32_t lane = lane_id();
    WaveOut o{NO_BANK, 0, 0};
    uint32_t idle = 0;
    unsigned long long idle_since = 0;  // lane 0: when this wave ran out of work
    if (lane == 0) wt = WaveTab{0, NO_PAGE, 0, 0, 0, 0};
    // per-wave counters in LDS (lane 0 updates them), flushed at exit
    __shared__ unsigned long long stats[ST_N];
    if (lane == 0)
        for (int i = 0; i < ST_N; i++) stats[i] = 0;
    unsigned long long t_wave = wall_clock64();
    unsigned long long starve_since = 0;
    unsigned long long deadline = 0, retire_at = ~0ull;  // retire_at: lane 0 only
    // a dormant wave has no first root chunk: the host starts root_cursor at
    // the number of waves that start at once
    const bool dormant = P.active_waves && blockIdx.x >= P.active_waves;
    bool first_claim = !dormant;  // lane 0 only
    if (lane == 0) {
        // Each wave times its deadline and slice from its own start (the
        // waves of a launch start within microseconds); only workgroup 0
        // publishes the launch start for the spin-wait deadline, so the
        // waves do not queue on one word.
        const unsigned long long now = wall_clock64();
        if (blockIdx.x == 0)
            __hip_atomic_store(&P.ctl->t0, now, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        deadline = now + P.timeout_ticks;
        if (P.serving && P.slice_ticks) retire_at = now + P.slice_ticks;
    }
    deadline = ((unsigned long long)bcast((uint32_t)(deadline >> 32), 0) << 32) | bcast((uint32_t)deadline, 0);
    if (dormant) {
        // Wait, polling every ~14 us, until the launch is join_ticks old
        // (then take part like any idle wave), has no work left, aborts or
        // passes its deadline (the loop below then leaves at once).
        if (lane == 0) {
            while (true) {
                for (int i = 0; i < 4; i++) __builtin_amdgcn_s_sleep(127);
                const unsigned long long now = wall_clock64();
                if (now - t_wave >= P.join_ticks || now > deadline || ld_rlx(&P.ctl->done) ||
                    ld_rlx(&P.ctl->abort))
                    break;
            }
        }
        __syncthreads();
        t_wave = wall_clock64();  // resident time counts from the join
    }
    while (true) {
        // 0: work, 1: idle, 2: exit
        if (lane == 0) {
            uint32_t flag = 1;
            if (ld_rlx(&P.ctl->abort)) {
                flag = 2;
            } else if (wall_clock64() > deadline) {
                timeout_at(P.ctl, 1, srv_active(ld_rlx(&P.ctl->srv_state)), (uint32_t)ld_rlx(&P.ctl->pending), idle);
                flag = 2;
            } else if (wall_clock64() > retire_at) {
                // serving: leave between tasks; queued tasks, admitted
                // queries and banks stay on the device for the next launch.
                // Closing stops admissions, so no query starts on the waves
                // still finishing their tasks.
                __hip_atomic_fetch_or(&P.ctl->srv_state, SRV_CLOSING, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                flag = 2;
            } else {
                TaskDesc t;
                const bool first = first_claim;
                first_claim = false;
                // A starting wave's own chunk (workgroup id) is claimed before
                // anything else: no other wave takes it (root_cursor starts
                // past the starting waves' chunks), so skipping it would leave
                // its roots unmatched and the launch waiting for them.
                if ((first && !P.serving && claim_root(P, t, stats, true)) || pop_task(P, t) ||
                    (!P.serving && claim_root(P, t, stats, false))) {
                    s_task = t;
                    flag = 0;
                } else if (P.serving) {
                    const uint32_t slot = claim_submission(P);
                    if (slot != NO_PAGE) {
                        s_task.q = slot;
                        flag = 3;
                    } else if (ld_sys(&P.srv->abort)) {
                        timeout_at(P.ctl, 3, srv_active(ld_rlx(&P.ctl->srv_state)), 0, 0);
                        flag = 2;
                    } else {
                        // The launch ends as a whole: a wave that has been
                        // idle long enough closes it, but only while no query
                        // is active and no submission waits (one CAS on the
                        // state word). Waves leave once it is closed, so a
                        // query is never admitted into a launch whose other
                        // waves have already left.
                        unsigned long long st = ld_rlx(&P.ctl->srv_state);
                        if (st & SRV_CLOSING) {
                            flag = 2;
                        } else if (srv_active(st) == 0 &&
                                   (ld_sys(&P.srv->shutdown) ||
                                    (idle_since && wall_clock64() - idle_since > P.idle_ticks)) &&
                                   ld_sys(&P.srv->sub_tail) == srv_head(st) &&
                                   __hip_atomic_compare_exchange_strong(&P.ctl->srv_state, &st, st | SRV_CLOSING,
                                                                        __ATOMIC_ACQ_REL, __ATOMIC_RELAXED,
                                                                        __HIP_MEMORY_SCOPE_AGENT)) {
                            flag = 2;
                        }
                    }
                } else if (ld_rlx(&P.ctl->done)) {
                    flag = 2;
                }
            }
            s_flag = flag;
        }
        __syncthreads();
        const uint32_t flag = s_flag;
        if (flag == 2) break;
        if (flag == 3) {
            const uint32_t slot = s_task.q;
            __syncthreads();
            const unsigned long long ta = wall_clock64();
            admit(P, slot, stats);
            if (lane == 0) stats[ST_BUSY] += wall_clock64() - ta;
            idle_since = 0;
            continue;
        }
        if (flag == 1) {
            if (idle == 0 && lane == 0) {
                atom_add(&P.ctl->idle_waves, 1);
                idle_since = wall_clock64();
            }
            // exponential back-off keeps idle waves off the queue counters
            for (uint32_t i = 0, lim = 1u << min(idle, (uint32_t)FM_IDLE_BACKOFF_LOG2); i < lim; i++)
                __builtin_amdgcn_s_sleep(8);
            idle++;
            continue;
        }
        if (idle && lane == 0) atom_add(&P.ctl->idle_waves, -1);
        idle = 0;
        const TaskDesc t = s_task;
        __syncthreads();
        if (P.instrument && lane == 0)
            for (uint32_t i = 0; i < 4 + MAXN; i++) s.inst[i] = 0;
        const unsigned long long tb = wall_clock64();
        // serving: a task whose tag is not the slot's current query belongs
        // to a query that has already completed; it is dropped (counted)
        uint32_t stale = 0;
        if (P.serving && lane == 0) stale = (t.q >> 24) != (ld_rlx(&P.qs[qslot(t.q)].pad) & 0xFFu);
        stale = bcast(stale, 0);
        if (!stale) process_task<PAIRED>(P, t, s, o, wt, stats, starve_since, deadline);
        else if (lane == 0) stats[ST_STALE]++;
        __syncthreads();
        if (lane == 0) stats[ST_BUSY] += wall_clock64() - tb;
        if (P.instrument && lane == 0) {
            DQuery &Q = P.qs[qslot(t.q)];
            unsigned long long *dst[4] = {&Q.row_words, &Q.lookups, &Q.candidates, &Q.probes};
            for (uint32_t i = 0; i < 4; i++)
                if (s.inst[i]) atom_add(dst[i], s.inst[i]);
            for (uint32_t i = 0; i < MAXN; i++)
                if (s.inst[4 + i]) atom_add(&Q.partials[i], s.inst[4 + i]);
        }
        if (lane == 0) {
            if (t.bank != ROOT_BANK && t.bank < P.nbanks) bank_release(P, t.bank);
            if (atom_add(&P.ctl->pending, -1ll) == 1 && !P.serving) st_rlx(&P.ctl->done, 1u);
            if (P.serving && !stale) {
                // rows of this task reach memory before the pending decrement
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                if (__hip_atomic_fetch_add(&P.qs[qslot(t.q)].qpending, -1ll, __ATOMIC_ACQ_REL,
                                           __HIP_MEMORY_SCOPE_AGENT) == 1)
                    complete_query(P, qslot(t.q));
            }
        }
    }
    if (lane == 0) {
        if (o.bank != NO_BANK) bank_release(P, o.bank);
        tab_release(P, wt);
        stats[ST_ALIVE] = wall_clock64() - t_wave;
    }
    __syncthreads();
    // one lane per counter, into this wave's shard
    if (lane < ST_N && stats[lane]) atom_add(&P.ctl->stats[blockIdx.x % STAT_SHARDS][lane], stats[lane]);
}

// One query (PAIRED: two candidates per lane), and serving (one per lane).
template <bool PAIRED>
__global__ __launch_bounds__(WAVE) void fm_match_kernel(KParams P) {
    match_waves<PAIRED>(P);
}

// Batches: two candidates per lane, compiled for 3 waves per SIMD (12 per
// CU), which spills a few registers; on 1,024 queries of 4-12 vertices
// 351-357 ms per launch against 362-366 ms with one candidate per lane at
// the same occupancy. One query runs the unspilled fm_match_kernel<true>
// instead (on R-MAT 2^20 with 16 queries of 6 vertices, one launch each:
// 67 ms of kernel time per step with this one against 64 ms).
__global__ __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(3, 3))) void fm_match_kernel_batch(KParams P) {
    match_waves<true>(P);
}

__global__ void fm_init_slots_kernel(Slot *ring, uint32_t qcap, uint32_t levels, FSlot *fring, uint32_t fcap) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t nq = (size_t)qcap * levels;
    if (i < nq) {
        Slot s{};
        s.seq = (uint32_t)(i & (qcap - 1));
        ring[i] = s;
    }
    if (i < fcap) {
        FSlot f{(uint32_t)i, 0};
        fring[i] = f;
    }
}

// ---------------------------------------------------------------------------
// Block-wide helpers of the .g parser kernels.
constexpr uint32_t IB = 256;  // threads per block

// Exclusive block scan over IB threads; total returned through *total.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *s_wave, uint32_t &total) {
    const uint32_t lane = lane_id(), wid = threadIdx.x / WAVE;
    const uint32_t incl = wave_incl_scan(v);
    if (lane == WAVE - 1) s_wave[wid] = incl;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
    for (uint32_t w = 0; w < IB / WAVE; w++) {
        if (w < wid) pre += s_wave[w];
        tot += s_wave[w];
    }
    __syncthreads();
    total = tot;
    return pre + incl - v;
}

// ---------------------------------------------------------------------------
// .g data file parser on the device (the reference parses on the GPU too:
// parseCSR, trie.cuh:800-951). The host checks the frame and header and
// uploads the body, lines of "v id label\n" / "e src dst label\n":
//  - fm_g_count_kernel: newlines per 4 KiB chunk (a 256-thread block, 16
//    bytes per thread); the host scans the chunk counts;
//  - fm_g_index_kernel: the start of every line;
//  - fm_g_parse_kernel: one thread per line, the host parser's checks; the
//    first malformed line (smallest index) is kept with an atomic minimum;
//  - fm_g_edges_kernel: the edge lines' values compacted in file order.
constexpr uint32_t GCHUNK = 4096;

// Bytes equal to '\n' in a 32-bit word (exact: no borrow between bytes).
__device__ __forceinline__ uint32_t newlines_in(uint32_t w) {
    const uint32_t x = w ^ 0x0A0A0A0Au;
    const uint32_t t = ~(((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x | 0x7F7F7F7Fu);
    return __popc(t);
}

__global__ __launch_bounds__(IB) void fm_g_count_kernel(const uint4 *body, uint64_t nchunks, uint32_t *counts) {
    __shared__ uint32_t s_wave[IB / WAVE];
    for (uint64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const uint4 v = body[c * IB + threadIdx.x];
        const uint32_t n = newlines_in(v.x) + newlines_in(v.y) + newlines_in(v.z) + newlines_in(v.w);
        uint32_t tot;
        block_excl_scan(n, s_wave, tot);
        if (threadIdx.x == 0) counts[c] = tot;
    }
}

__global__ __launch_bounds__(IB) void fm_g_index_kernel(const uint4 *body, uint64_t nchunks, const uint64_t *chunk_off,
                                                        uint64_t *starts) {
    __shared__ uint32_t s_wave[IB / WAVE];
    for (uint64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const uint4 v = body[c * IB + threadIdx.x];
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
        uint32_t n = 0;
#pragma unroll
        for (int i = 0; i < 4; i++) n += newlines_in(w[i]);
        uint32_t tot;
        uint64_t k = chunk_off[c] + block_excl_scan(n, s_wave, tot);
        const uint64_t pos0 = (c * IB + threadIdx.x) * 16;
#pragma unroll
        for (int i = 0; i < 16; i++)
            if (((w[i >> 2] >> (8 * (i & 3))) & 0xFFu) == '\n') starts[1 + k++] = pos0 + i + 1;
    }
}

__device__ __forceinline__ bool g_parse_uint(const char *body, uint64_t &p, uint64_t end, uint32_t &v) {
    if (p >= end || body[p] < '0' || body[p] > '9') return false;
    unsigned long long x = 0;
    while (p < end && body[p] >= '0' && body[p] <= '9') {
        x = x * 10 + (unsigned long long)(body[p] - '0');
        if (x > 0xFFFFFFFFull) return false;
        p++;
    }
    v = (uint32_t)x;
    return true;
}

// Line i is body[starts[i], starts[i + 1] - 1) followed by its newline.
__global__ __launch_bounds__(IB) void fm_g_parse_kernel(const char *body, const uint64_t *starts, uint64_t nlines,
                                                        uint32_t V, uint8_t *is_edge, uint32_t *vals,
                                                        unsigned long long *first_bad) {
    for (uint64_t i = (uint64_t)blockIdx.x * IB + threadIdx.x; i < nlines; i += (uint64_t)gridDim.x * IB) {
        uint64_t p = starts[i];
        const uint64_t end = starts[i + 1] - 1;  // the newline
        const char kind = p < end ? body[p] : '\n';
        bool ok = (kind == 'v' || kind == 'e') && p + 1 < end && body[p + 1] == ' ';
        uint32_t a = 0, b = 0, c = 0;
        if (ok) {
            p += 2;
            ok = g_parse_uint(body, p, end, a) && p < end && body[p++] == ' ' && g_parse_uint(body, p, end, b);
            if (ok && kind == 'e') ok = p < end && body[p++] == ' ' && g_parse_uint(body, p, end, c);
            ok = ok && p == end;
            if (ok && kind == 'v') ok = a < V;
        }
        if (!ok) {
            atomicMin(first_bad, (unsigned long long)i);
            is_edge[i] = 0;
            continue;
        }
        is_edge[i] = kind == 'e' ? 1 : 0;
        vals[3 * i] = a;
        vals[3 * i + 1] = b;
        vals[3 * i + 2] = c;
    }
}

// Edge lines per block of IB lines (counts), then their values scattered to
// src/dst/lab at the block's offset (edge_off) in line order.
__global__ __launch_bounds__(IB) void fm_g_edges_kernel(const uint8_t *is_edge, const uint32_t *vals, uint64_t nlines,
                                                        uint64_t nblocks, const uint64_t *edge_off, uint32_t *counts,
                                                        uint32_t *src, uint32_t *dst, uint32_t *lab) {
    __shared__ uint32_t s_wave[IB / WAVE];
    for (uint64_t blk = blockIdx.x; blk < nblocks; blk += gridDim.x) {
        const uint64_t i = blk * IB + threadIdx.x;
        const uint32_t e = i < nlines ? is_edge[i] : 0u;
        uint32_t tot;
        const uint32_t pre = block_excl_scan(e, s_wave, tot);
        if (!edge_off) {
            if (threadIdx.x == 0) counts[blk] = tot;
        } else if (e) {
            const uint64_t k = edge_off[blk] + pre;
            src[k] = vals[3 * i];
            dst[k] = vals[3 * i + 1];
            lab[k] = vals[3 * i + 2];
        }
    }
}

// Symmetry of a graph's out-rows: every edge (u, v) of label l has the edge
// (v, u) of label l. A wave takes 64 consecutive rows (row l * V + u), keeps
// their starts in LDS, steps over their edges 64 at a time (a hub row is
// spread over the lanes like any other) and probes each edge's reverse row.
// A missing reverse edge sets *asym; waves stop at their next chunk once it
// is set.
__global__ __launch_bounds__(WAVE) void fm_sym_check_kernel(const uint32_t *ptr, const uint32_t *col,
                                                          const uint32_t *skip, unsigned long long rows, uint32_t V,
                                                          uint32_t *asym) {
    __shared__ uint32_t rs[WAVE + 1];           // row starts of the chunk, and its end
    __shared__ uint32_t rlab[WAVE], rsrc[WAVE];  // label id and source vertex of each row
    const uint32_t lane = threadIdx.x;
    const unsigned long long nchunks = (rows + WAVE - 1) / WAVE;
    for (unsigned long long c = blockIdx.x; c < nchunks; c += gridDim.x) {
        if (__builtin_amdgcn_readfirstlane(__hip_atomic_load(asym, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) break;
        const unsigned long long r0 = c * WAVE, r = r0 + lane;
        // rows past the last one start at the end of the column array
        rs[lane] = ptr[min(r, rows)];
        if (lane == 0) rs[WAVE] = ptr[min(r0 + WAVE, rows)];
        const unsigned long long rc = min(r, rows - 1);
        rlab[lane] = (uint32_t)(rc / V);
        rsrc[lane] = (uint32_t)(rc % V);
        __syncthreads();
        const uint32_t e1 = rs[WAVE];
        for (uint32_t base = rs[0]; base < e1; base += WAVE) {
            const uint32_t e = base + lane;
            if (e < e1) {
                // the edge's row: the last j with rs[j] <= e (empty rows
                // share their start with the next row, so this is the
                // non-empty row holding e)
                uint32_t lo = 0, hi = WAVE - 1;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi + 1) >> 1;
                    if (rs[mid] <= e) lo = mid;
                    else hi = mid - 1;
                }
                const uint32_t v = col[e];
                const unsigned long long rr = (unsigned long long)rlab[lo] * V + v;
                const uint32_t st = ptr[rr], len = ptr[rr + 1] - st;
                if (!row_contains(col, skip, st, len, rsrc[lo]))
                    __hip_atomic_store(asym, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        __syncthreads();  // the next chunk rewrites the row starts
    }
}

}  // namespace fm

// ===========================================================================
// Host side.
using namespace fm;

#define HIP_TRY(expr)                                                                       \
    do {                                                                                    \
        hipError_t _e = (expr);                                                             \
        if (_e != hipSuccess)                                                               \
            return fail(_e == hipErrorOutOfMemory ? FM_ERR_OUT_OF_MEMORY : FM_ERR_DEVICE,   \
                        "%s failed: %s", #expr, hipGetErrorString(_e));                     \
    } while (0)

struct fm_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int cu_count = 0;
    int waves_per_cu = 0;         // resident waves per CU: fm_match_kernel<false>
    int waves_per_cu_paired = 0;  // the same for fm_match_kernel<true>
    int waves_per_cu_batch = 0;   // and for fm_match_kernel_batch
    uint32_t grid_waves = 0;
    double wall_hz = 100e6;
    char name[64] = {0};
    // arenas
    uint32_t *banks = nullptr;
    uint64_t bank_words = 0;
    uint32_t nbanks = 0;
    int32_t *bank_ref = nullptr;
    FSlot *free_ring = nullptr;
    uint32_t fcap = 0;
    Slot *ring = nullptr;
    uint32_t qcap = 0;
    uint32_t *pages = nullptr;
    uint64_t page_words = 0;
    uint32_t npages = 0;
    uint32_t *page_owner = nullptr;
    uint32_t *page_fill = nullptr;
    Ctl *ctl = nullptr;
    // per-launch buffers (grow only)
    DPlan *d_plans = nullptr;
    DQuery *d_qs = nullptr;
    size_t q_cap = 0;
    uint32_t max_queries = 0;
    uint64_t epoch = 0;
    bool dirty = true;  // queues/pool need a full re-initialisation
    fm_server *server = nullptr;  // a server owns the engine while it lives
    void *isect_buf = nullptr;    // intersection work array (regrown when too small, one layout each)
    uint64_t isect_segs = 0, isect_pairs = 0;  // its segment and pair capacity
    FSlot *pfree_ring = nullptr;
    uint32_t pfcap = 0;
    uint64_t prealloc = 0;     // PREALLOC: arenas + graph may take this many device bytes
    uint64_t arena_bytes = 0;  // device bytes of the engine's arenas
};

// PREALLOC (src/environment.cuh:4): the reference bump-allocates the graph
// and its arenas from one slab and exits with status 1 when they do not fit
// (src/trie.cuh:1034-1049). Here the arenas are separate allocations; the
// same budget is checked before a launch on graph g.
static int check_prealloc(const fm_engine *e, const fm_graph *g) {
    const uint64_t need = e->arena_bytes + (g ? g->device_bytes : 0);
    if (e->prealloc != FM_PREALLOC_UNLIMITED && need > e->prealloc)
        return fail(FM_ERR_GENERIC, "PREALLOC exceeded: arenas %llu bytes + graph %llu bytes > PREALLOC %llu bytes",
                    (unsigned long long)e->arena_bytes, (unsigned long long)(g ? g->device_bytes : 0),
                    (unsigned long long)e->prealloc);
    return FM_OK;
}

// Launch deadline in wall-clock ticks; FM_NO_DEADLINE gives one no launch
// reaches (2^62 ticks), so deadline arithmetic never wraps.
static unsigned long long deadline_ticks(uint32_t timeout_ms, double wall_hz) {
    if (timeout_ms == FM_NO_DEADLINE) return 1ull << 62;
    return (unsigned long long)((double)timeout_ms * (wall_hz / 1000.0));
}

static int isect_scratch_init(int device);

// A launch counter summed over its shards.
static unsigned long long stat_sum(const Ctl &c, uint32_t i) {
    unsigned long long v = 0;
    for (uint32_t k = 0; k < STAT_SHARDS; k++) v += c.stats[k][i];
    return v;
}

struct fm_result {
    fm_engine *e = nullptr;
    uint64_t epoch = 0;
    std::vector<DQuery> qs;
    std::vector<uint32_t> n;
    fm_launch_stats stats{};
    uint32_t pages_used = 0;
    std::vector<uint32_t> page_owner, page_fill;  // fetched on the first table copy
};

static int full_reset(fm_engine *e) {
    HIP_TRY(hipSetDevice(e->device));
    HIP_TRY(hipMemsetAsync(e->ctl, 0, sizeof(Ctl), e->stream));
    size_t total = std::max<size_t>((size_t)e->qcap * NLEVELS, e->fcap);
    unsigned blocks = (unsigned)((total + 255) / 256);
    hipLaunchKernelGGL(fm_init_slots_kernel, dim3(blocks), dim3(256), 0, e->stream, e->ring, e->qcap,
                       NLEVELS, e->free_ring, e->fcap);
    HIP_TRY(hipGetLastError());
    if (e->pfree_ring) {
        hipLaunchKernelGGL(fm_init_slots_kernel, dim3((e->pfcap + 255) / 256), dim3(256), 0, e->stream, e->ring, 0u,
                           0u, e->pfree_ring, e->pfcap);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->dirty = false;
    return FM_OK;
}

// Where a launch ran out of time, from Ctl::diag.
static std::string timeout_site(const Ctl &c) {
    static const char *names[] = {"?", "main loop", "task", "host abort", "?", "?", "?", "?", "?", "?", "?",
                                  "bank from free ring", "bank release", "task push", "task pop",
                                  "page from free ring", "?", "page recycle", "serving table page"};
    const uint32_t site = c.diag[0];
    char buf[160];
    snprintf(buf, sizeof buf, "at %s, values %u %u %u", site < sizeof(names) / sizeof(names[0]) ? names[site] : "?",
             c.diag[1], c.diag[2], c.diag[3]);
    return buf;
}

static uint32_t pow2_at_least(uint64_t x) {
    uint32_t p = 1;
    while (p < x && p < (1u << 31)) p <<= 1;
    return p;
}

extern "C" {

int fm_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int fm_graph_upload(fm_graph *g) {
    if (!g) return fail(FM_ERR_INVALID_ARG, "null graph");
    if (g->device >= 0) return FM_OK;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    const HostGraph &h = g->h;
    uint64_t E = h.E();
    if (!h.symmetric && 2 * E >= (1ull << 32))
        return fail(FM_ERR_NOT_SUPPORTED, "directed graph with %llu edges exceeds the u32 row index",
                    (unsigned long long)E);
    uint64_t rows = (uint64_t)h.L() * h.V + 1;
    uint64_t bytes = 0;
    auto dalloc = [&](uint32_t **p, uint64_t n) -> int {
        HIP_TRY(hipMalloc((void **)p, std::max<uint64_t>(n, 1) * 4));
        bytes += n * 4;
        return FM_OK;
    };
    int rc;
    if ((rc = dalloc(&g->d_ptr[0], rows))) return rc;
    HIP_TRY(hipMemcpy(g->d_ptr[0], h.out.ptr.data(), rows * 4, hipMemcpyHostToDevice));
    uint64_t ncol = h.symmetric ? E : 2 * E;
    if ((rc = dalloc(&g->d_col[0], ncol + COL_PAD))) return rc;
    HIP_TRY(hipMemcpy(g->d_col[0], h.out.col.data(), E * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(g->d_col[0] + ncol, 0, COL_PAD * 4));
    {
        // skip index over the device column array (out-rows, then in-rows)
        const uint64_t nskip = (ncol + (1u << SKIP_SHIFT) - 1) >> SKIP_SHIFT;
        std::vector<uint32_t> skip(nskip + COL_PAD, 0u);
        for (uint64_t b = 0; b < nskip; b++) {
            const uint64_t i = b << SKIP_SHIFT;
            skip[b] = i < E ? h.out.col[i] : h.in.col[i - E];
        }
        if ((rc = dalloc(&g->d_skip, skip.size()))) return rc;
        HIP_TRY(hipMemcpy(g->d_skip, skip.data(), skip.size() * 4, hipMemcpyHostToDevice));
    }
    uint64_t nroots = h.out.roots.size() + (h.symmetric ? 0 : h.in.roots.size());
    if ((rc = dalloc(&g->d_roots[0], nroots))) return rc;
    HIP_TRY(hipMemcpy(g->d_roots[0], h.out.roots.data(), h.out.roots.size() * 4, hipMemcpyHostToDevice));
    if (h.symmetric) {
        g->d_ptr[1] = g->d_ptr[0];
        g->d_col[1] = g->d_col[0];
        g->d_roots[1] = g->d_roots[0];
    } else {
        std::vector<uint32_t> shifted(h.in.ptr);
        for (auto &v : shifted) v += (uint32_t)E;
        if ((rc = dalloc(&g->d_ptr[1], rows))) return rc;
        HIP_TRY(hipMemcpy(g->d_ptr[1], shifted.data(), rows * 4, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(g->d_col[0] + E, h.in.col.data(), E * 4, hipMemcpyHostToDevice));
        g->d_col[1] = g->d_col[0];
        HIP_TRY(hipMemcpy(g->d_roots[0] + h.out.roots.size(), h.in.roots.data(), h.in.roots.size() * 4,
                          hipMemcpyHostToDevice));
        g->d_roots[1] = g->d_roots[0];
    }
    g->device = dev;
    g->device_bytes = bytes;
    return FM_OK;
}

}  // extern "C"

// Parses the body of a .g data file on the current device into edge lists
// (in file order). Errors carry the host parser's codes and messages: the
// first malformed line found on the device is parsed again on the host for
// the message.
static int gpu_parse_data(const std::string &buf, const DataFrame &f, uint32_t *&d_edges, uint64_t &n_edges) {
    d_edges = nullptr;
    n_edges = 0;
    const uint64_t nbody = f.body_end - f.body_begin;
    if (nbody == 0) return FM_OK;
    const uint64_t nchunks = (nbody + GCHUNK - 1) / GCHUNK;
    std::vector<void *> owned;
    struct Free {
        std::vector<void *> &v;
        ~Free() {
            for (void *p : v) (void)hipFree(p);
        }
    } guard{owned};
    auto dalloc = [&](void **p, uint64_t bytes) -> int {
        HIP_TRY(hipMalloc(p, std::max<uint64_t>(bytes, 16)));
        owned.push_back(*p);
        return FM_OK;
    };
    int rc;
    char *d_body = nullptr;
    if ((rc = dalloc((void **)&d_body, nchunks * GCHUNK))) return rc;
    HIP_TRY(hipMemset(d_body + nbody, 0, nchunks * GCHUNK - nbody));
    HIP_TRY(hipMemcpy(d_body, buf.data() + f.body_begin, nbody, hipMemcpyHostToDevice));
    const unsigned grid = (unsigned)std::min<uint64_t>(nchunks, 65536);
    // newlines per chunk, their offsets (host scan), the line starts
    uint32_t *d_counts = nullptr;
    if ((rc = dalloc((void **)&d_counts, nchunks * 4))) return rc;
    hipLaunchKernelGGL(fm_g_count_kernel, dim3(grid), dim3(IB), 0, 0, (const uint4 *)d_body, nchunks, d_counts);
    HIP_TRY(hipGetLastError());
    std::vector<uint32_t> counts(nchunks);
    HIP_TRY(hipMemcpy(counts.data(), d_counts, nchunks * 4, hipMemcpyDeviceToHost));
    std::vector<uint64_t> off(nchunks);
    uint64_t nl = 0;
    for (uint64_t c = 0; c < nchunks; c++) {
        off[c] = nl;
        nl += counts[c];
    }
    // bytes after the last newline form an unterminated (malformed) line
    const bool tail = buf[f.body_end - 1] != '\n';
    uint64_t *d_off = nullptr, *d_starts = nullptr;
    if ((rc = dalloc((void **)&d_off, nchunks * 8))) return rc;
    if ((rc = dalloc((void **)&d_starts, (nl + 1) * 8))) return rc;
    HIP_TRY(hipMemcpy(d_off, off.data(), nchunks * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(d_starts, 0, 8));
    hipLaunchKernelGGL(fm_g_index_kernel, dim3(grid), dim3(IB), 0, 0, (const uint4 *)d_body, nchunks,
                       (const uint64_t *)d_off, d_starts);
    HIP_TRY(hipGetLastError());
    // parse every complete line
    uint8_t *d_is_edge = nullptr;
    uint32_t *d_vals = nullptr;
    unsigned long long *d_bad = nullptr;
    if ((rc = dalloc((void **)&d_is_edge, nl))) return rc;
    if ((rc = dalloc((void **)&d_vals, nl * 12))) return rc;
    if ((rc = dalloc((void **)&d_bad, 8))) return rc;
    HIP_TRY(hipMemset(d_bad, 0xFF, 8));
    const uint64_t nblocks = (nl + IB - 1) / IB;
    const unsigned lgrid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(nblocks, 65536));
    if (nl) {
        hipLaunchKernelGGL(fm_g_parse_kernel, dim3(lgrid), dim3(IB), 0, 0, (const char *)d_body,
                           (const uint64_t *)d_starts, nl, f.V, d_is_edge, d_vals, d_bad);
        HIP_TRY(hipGetLastError());
    }
    unsigned long long bad = ~0ull;
    HIP_TRY(hipMemcpy(&bad, d_bad, 8, hipMemcpyDeviceToHost));
    if (bad != ~0ull || tail) {
        // the first bad line: its start is the end of the line before
        uint64_t i = bad != ~0ull ? bad : nl, start = 0;
        if (i > 0) HIP_TRY(hipMemcpy(&start, d_starts + i, 8, hipMemcpyDeviceToHost));
        const char *p = buf.data() + f.body_begin + start, *stop = buf.data() + f.body_end;
        char kind;
        uint32_t vals[3];
        rc = parse_data_line(p, stop, f.V, i + 3, kind, vals);
        return rc ? rc : fail(FM_ERR_INTERNAL, "device parser rejected line %llu", (unsigned long long)(i + 3));
    }
    // edge lines, compacted in file order
    uint32_t *d_ecount = nullptr;
    uint64_t *d_eoff = nullptr;
    if ((rc = dalloc((void **)&d_ecount, nblocks * 4))) return rc;
    if ((rc = dalloc((void **)&d_eoff, nblocks * 8))) return rc;
    if (nl) {
        hipLaunchKernelGGL(fm_g_edges_kernel, dim3(lgrid), dim3(IB), 0, 0, (const uint8_t *)d_is_edge,
                           (const uint32_t *)d_vals, nl, nblocks, (const uint64_t *)nullptr, d_ecount, nullptr,
                           nullptr, nullptr);
        HIP_TRY(hipGetLastError());
    }
    std::vector<uint32_t> ecount(nblocks);
    if (nblocks) HIP_TRY(hipMemcpy(ecount.data(), d_ecount, nblocks * 4, hipMemcpyDeviceToHost));
    std::vector<uint64_t> eoff(nblocks);
    uint64_t ne = 0;
    for (uint64_t b = 0; b < nblocks; b++) {
        eoff[b] = ne;
        ne += ecount[b];
    }
    uint32_t *d_e = nullptr;
    if ((rc = dalloc((void **)&d_e, ne * 12))) return rc;
    if (nblocks) HIP_TRY(hipMemcpy(d_eoff, eoff.data(), nblocks * 8, hipMemcpyHostToDevice));
    if (ne) {
        hipLaunchKernelGGL(fm_g_edges_kernel, dim3(lgrid), dim3(IB), 0, 0, (const uint8_t *)d_is_edge,
                           (const uint32_t *)d_vals, nl, nblocks, (const uint64_t *)d_eoff, d_ecount, d_e,
                           d_e + ne, d_e + 2 * ne);
        HIP_TRY(hipGetLastError());
    }
    // the edge arrays stay on the device, owned by the caller
    owned.erase(std::find(owned.begin(), owned.end(), (void *)d_e));
    d_edges = d_e;
    n_edges = ne;
    return FM_OK;
}

// Copies device edge arrays src[n] | dst[n] | label[n] to the host.
static int edges_to_host(const uint32_t *d_e, uint64_t n, std::vector<uint32_t> &src, std::vector<uint32_t> &dst,
                         std::vector<uint32_t> &lab) {
    src.resize(n);
    dst.resize(n);
    lab.resize(n);
    if (n) {
        HIP_TRY(hipMemcpy(src.data(), d_e, n * 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(dst.data(), d_e + n, n * 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(lab.data(), d_e + 2 * n, n * 4, hipMemcpyDeviceToHost));
    }
    return FM_OK;
}

extern "C" {

}  // extern "C"

// Frees a graph's device arrays (the host CSR stays).
static void graph_free_device(fm_graph *g) {
    if (g->device < 0) return;
    (void)hipFree(g->d_ptr[0]);
    if (g->d_ptr[1] != g->d_ptr[0]) (void)hipFree(g->d_ptr[1]);
    (void)hipFree(g->d_col[0]);
    (void)hipFree(g->d_skip);
    (void)hipFree(g->d_roots[0]);
    g->d_ptr[0] = g->d_ptr[1] = g->d_col[0] = g->d_col[1] = g->d_skip = g->d_roots[0] = g->d_roots[1] = nullptr;
    g->device = -1;
    g->device_bytes = 0;
}

// Whether every out-edge (u, v, l) of an uploaded graph has its reverse
// (v, u, l), checked on the device (fm_sym_check_kernel).
static int gpu_symmetric(const fm_graph *g, bool &sym) {
    const HostGraph &h = g->h;
    sym = true;
    const uint64_t rows = (uint64_t)h.L() * h.V;
    if (h.E() == 0 || rows == 0) return FM_OK;
    uint32_t *d_flag = nullptr;
    HIP_TRY(hipMalloc((void **)&d_flag, 4));
    const uint64_t nchunks = (rows + WAVE - 1) / WAVE;
    const uint32_t grid = (uint32_t)std::min<uint64_t>(nchunks, 8192);
    hipError_t err = hipMemset(d_flag, 0, 4);
    if (err == hipSuccess) {
        hipLaunchKernelGGL(fm_sym_check_kernel, dim3(grid), dim3(WAVE), 0, 0, g->d_ptr[0], g->d_col[0], g->d_skip,
                           (unsigned long long)rows, h.V, d_flag);
        err = hipGetLastError();
    }
    uint32_t flag = 0;
    if (err == hipSuccess) err = hipMemcpy(&flag, d_flag, 4, hipMemcpyDeviceToHost);
    (void)hipFree(d_flag);
    if (err != hipSuccess) return fail(FM_ERR_DEVICE, "symmetry check: %s", hipGetErrorString(err));
    sym = flag == 0;
    return FM_OK;
}

extern "C" {

int fm_graph_load_g(const char *path, fm_graph **out) {
    if (!path || !out) return fail(FM_ERR_INVALID_ARG, "null argument");
    std::string buf;
    if (!read_file(path, buf)) return fail(FM_ERR_DATA_MISSING_FILE, "missing file %s", path);
    DataFrame f;
    int rc = parse_data_frame(buf, f);
    if (rc) return rc;
    uint32_t *d_edges = nullptr;
    uint64_t ne = 0;
    if ((rc = gpu_parse_data(buf, f, d_edges, ne))) return rc;
    struct EdgesFree {
        uint32_t *p;
        ~EdgesFree() { (void)hipFree(p); }
    } edges_guard{d_edges};
    buf.clear();
    buf.shrink_to_fit();
    auto *g = new fm_graph();
    g->h.header_edges = f.header_edges;
    // The out-rows are built on the device as if every line had its reverse
    // (on the host when gpu_build_graph declines), and the symmetry is
    // checked on the device. A directed graph is then built again on the
    // host with both directions.
    bool built = false;
    rc = gpu_build_graph(f.V, ne, d_edges, g, built);
    if (!rc && !built) {
        std::vector<uint32_t> src, dst, lab;
        rc = edges_to_host(d_edges, ne, src, dst, lab);
        if (!rc) rc = build_graph(f.V, ne, src.data(), dst.data(), lab.data(), g->h, true);
        if (!rc) rc = fm_graph_upload(g);
    }
    bool sym = true;
    if (!rc) rc = gpu_symmetric(g, sym);
    if (!rc && !sym) {
        graph_free_device(g);
        g->h = HostGraph();
        g->h.header_edges = f.header_edges;
        std::vector<uint32_t> src, dst, lab;
        rc = edges_to_host(d_edges, ne, src, dst, lab);
        if (!rc) rc = build_graph(f.V, ne, src.data(), dst.data(), lab.data(), g->h);
        if (!rc) rc = fm_graph_upload(g);
    }
    if (rc) {
        fm_graph_destroy(g);
        return rc;
    }
    *out = g;
    return FM_OK;
}

int fm_graph_copy_device_csr(const fm_graph *g, int dir, uint32_t *ptr, uint32_t *col) {
    if (!g || g->device < 0 || (dir != 0 && dir != 1)) return fail(FM_ERR_INVALID_ARG, "graph not uploaded");
    const HostGraph &h = g->h;
    uint64_t rows = (uint64_t)h.L() * h.V + 1, E = h.E();
    bool in = dir == 1 && !h.symmetric;
    if (ptr) {
        HIP_TRY(hipMemcpy(ptr, g->d_ptr[dir], rows * 4, hipMemcpyDeviceToHost));
        if (in)
            for (uint64_t i = 0; i < rows; i++) ptr[i] -= (uint32_t)E;
    }
    if (col) HIP_TRY(hipMemcpy(col, g->d_col[0] + (in ? E : 0), E * 4, hipMemcpyDeviceToHost));
    return FM_OK;
}

void fm_graph_destroy(fm_graph *g) {
    if (!g) return;
    graph_free_device(g);
    delete g;
}

int fm_engine_create(const fm_engine_config *cfg, fm_engine **out) {
    if (!out) return fail(FM_ERR_INVALID_ARG, "null argument");
    fm_engine_config c{};
    if (cfg) c = *cfg;
    if (!c.soln_prealloc) c.soln_prealloc = SOLN_PREALLOC;
    if (!c.output_prealloc) c.output_prealloc = OUTPUT_PREALLOC;
    if (!c.bank_size) c.bank_size = CurBankSize;
    if (!c.output_page_bytes) c.output_page_bytes = OUTPUT_PAGE_BYTES;
    if (!c.queue_capacity) c.queue_capacity = 1u << 20;
    if (!c.max_queries) c.max_queries = 1u << 16;
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (ndev == 0) return fail(FM_ERR_DEVICE, "no HIP device");
    if (c.device < 0 || c.device >= ndev) return fail(FM_ERR_INVALID_ARG, "device %d out of range", c.device);
    auto *e = new fm_engine();
    e->device = c.device;
    e->max_queries = c.max_queries;
    auto cleanup_fail = [&](int rc) {
        fm_engine_destroy(e);
        return rc;
    };
    if (hipSetDevice(e->device) != hipSuccess) return cleanup_fail(fail(FM_ERR_DEVICE, "hipSetDevice failed"));
    if (int rc = isect_scratch_init(e->device)) return cleanup_fail(rc);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, e->device) != hipSuccess)
        return cleanup_fail(fail(FM_ERR_DEVICE, "hipGetDeviceProperties failed"));
    e->cu_count = prop.multiProcessorCount;
    snprintf(e->name, sizeof e->name, "%s", prop.gcnArchName);
    int wall_khz = 0;
    if (hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, e->device) == hipSuccess && wall_khz > 0)
        e->wall_hz = wall_khz * 1000.0;
    int blocks = 0, blocks2 = 0, blocks3 = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, fm_match_kernel<false>, WAVE, 0) != hipSuccess ||
        blocks <= 0)
        blocks = 8;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks2, fm_match_kernel<true>, WAVE, 0) != hipSuccess ||
        blocks2 <= 0)
        blocks2 = std::min(blocks, 8);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks3, fm_match_kernel_batch, WAVE, 0) != hipSuccess ||
        blocks3 <= 0)
        blocks3 = std::min(blocks, 8);
    e->waves_per_cu = blocks;
    e->waves_per_cu_paired = blocks2;
    e->waves_per_cu_batch = blocks3;
    // Default grid: 8 waves per CU. The serving launch uses it as is; a
    // one-query launch starts half of it and a batch uses every resident
    // wave (fm_match_batch gives the measurements).
    e->grid_waves = (uint32_t)(std::min(blocks, 8) * e->cu_count);
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&e->ev0) != hipSuccess || hipEventCreate(&e->ev1) != hipSuccess)
        return cleanup_fail(fail(FM_ERR_DEVICE, "stream/event creation failed"));
    // A fresh bank must hold one candidate step's children: up to 2 x 64
    // rows (two candidates per lane) of up to MAXN words, i.e. WAVE * MAXN
    // 8-byte units; a bank that already holds rows is flushed before a step
    // that would not fit.
    if (c.bank_size < MIN_BANK_SIZE)
        return cleanup_fail(fail(FM_ERR_INVALID_ARG, "bank_size %u is below the minimum %u (one step of rows)",
                                 c.bank_size, MIN_BANK_SIZE));
    e->bank_words = (uint64_t)c.bank_size * 2;  // 8-byte units -> u32 words
    e->nbanks = (uint32_t)std::min<uint64_t>(c.soln_prealloc / (e->bank_words * 4), 1u << 30);
    if (e->nbanks == 0) return cleanup_fail(fail(FM_ERR_INVALID_ARG, "SOLN_PREALLOC smaller than one bank"));
    e->fcap = pow2_at_least(e->nbanks);
    e->qcap = pow2_at_least(c.queue_capacity);
    e->page_words = c.output_page_bytes / 4;
    e->npages = (uint32_t)std::min<uint64_t>(c.output_prealloc / c.output_page_bytes, 1u << 30);
    if (e->npages == 0) return cleanup_fail(fail(FM_ERR_INVALID_ARG, "OUTPUT_PREALLOC smaller than one page"));
    if (e->page_words < WAVE * MAXN)
        return cleanup_fail(fail(FM_ERR_INVALID_ARG, "output pages must hold at least %u words", WAVE * MAXN));
    e->prealloc = c.prealloc ? c.prealloc : (uint64_t)PREALLOC;
    e->arena_bytes = (uint64_t)e->nbanks * e->bank_words * 4 + (uint64_t)e->nbanks * 4 +
                     (uint64_t)e->fcap * sizeof(FSlot) + (uint64_t)e->qcap * NLEVELS * sizeof(Slot) +
                     (uint64_t)e->npages * e->page_words * 4 + (uint64_t)e->npages * 8 + sizeof(Ctl);
    if (int rc = check_prealloc(e, nullptr)) return cleanup_fail(rc);
    auto alloc = [&](void **p, size_t bytes) { return hipMalloc(p, std::max<size_t>(bytes, 4)) == hipSuccess; };
    if (!alloc((void **)&e->banks, (size_t)e->nbanks * e->bank_words * 4) ||
        !alloc((void **)&e->bank_ref, (size_t)e->nbanks * 4) ||
        !alloc((void **)&e->free_ring, (size_t)e->fcap * sizeof(FSlot)) ||
        !alloc((void **)&e->ring, (size_t)e->qcap * NLEVELS * sizeof(Slot)) ||
        !alloc((void **)&e->pages, (size_t)e->npages * e->page_words * 4) ||
        !alloc((void **)&e->page_owner, (size_t)e->npages * 4) || !alloc((void **)&e->page_fill, (size_t)e->npages * 4) ||
        !alloc((void **)&e->ctl, sizeof(Ctl)))
        return cleanup_fail(fail(FM_ERR_OUT_OF_MEMORY, "engine arenas do not fit in device memory"));
    int rc = full_reset(e);
    if (rc) return cleanup_fail(rc);
    *out = e;
    return FM_OK;
}

void fm_engine_destroy(fm_engine *e) {
    if (!e) return;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    (void)hipFree(e->banks);
    (void)hipFree(e->bank_ref);
    (void)hipFree(e->free_ring);
    (void)hipFree(e->isect_buf);
    (void)hipFree(e->ring);
    (void)hipFree(e->pages);
    (void)hipFree(e->ctl);
    (void)hipFree(e->page_owner);
    (void)hipFree(e->page_fill);
    (void)hipFree(e->pfree_ring);
    (void)hipFree(e->d_plans);
    (void)hipFree(e->d_qs);
    if (e->ev0) (void)hipEventDestroy(e->ev0);
    if (e->ev1) (void)hipEventDestroy(e->ev1);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

int fm_engine_info_get(fm_engine *e, fm_engine_info *out) {
    if (!e || !out) return fail(FM_ERR_INVALID_ARG, "null argument");
    out->compute_units = (uint32_t)e->cu_count;
    out->waves_per_cu = (uint32_t)e->waves_per_cu;
    out->grid_waves = e->grid_waves;
    out->banks = e->nbanks;
    out->bank_bytes = e->bank_words * 4;
    out->output_pages = e->npages;
    out->output_page_bytes = (uint32_t)(e->page_words * 4);
    snprintf(out->device_name, sizeof out->device_name, "%s", e->name);
    out->arena_bytes = e->arena_bytes;
    return FM_OK;
}

// Device plan of a query against a graph: dense label ids, symmetric
// de-duplication, root list selection.
static int make_dplan(const fm_graph *g, const HostQuery &q, bool all_roots, uint32_t plan, uint32_t symmetry,
                      DPlan &dp) {
    const HostGraph &h = g->h;
    if (plan != FM_PLAN_COST && plan != FM_PLAN_REFERENCE) return fail(FM_ERR_INVALID_ARG, "unknown plan %u", plan);
    Plan cost_plan;
    if (plan == FM_PLAN_COST) {
        int rc = make_cost_plan(q, h, cost_plan);
        if (rc) return rc;
    }
    const Plan &p = plan == FM_PLAN_COST ? cost_plan : q.plan;
    std::memset(&dp, 0, sizeof dp);
    dp.n = p.n;
    uint32_t nr = 0;
    for (uint32_t d = 0; d < p.n; d++) {
        dp.order[d] = p.order[d];
        dp.req_begin[d] = nr;
        std::vector<uint32_t> anchored, self;
        for (uint32_t i = p.req_begin[d]; i < p.req_begin[d + 1]; i++) {
            const Req &r = p.reqs[i];
            uint32_t lid = h.label_id(r.label);
            if (lid == UINT32_MAX) lid = LABEL_ABSENT;
            uint32_t dir = h.symmetric ? 0 : r.dir;
            uint32_t packed = (r.anchor & 63u) | (dir << 7) | (lid << 8);
            auto &dst = r.anchor == d ? self : anchored;
            if (std::find(dst.begin(), dst.end(), packed) == dst.end()) dst.push_back(packed);
        }
        for (auto v : anchored) dp.req[nr++] = v;
        for (auto v : self) dp.req[nr++] = v;
        dp.nself[d] = (uint32_t)self.size();
    }
    dp.req_begin[p.n] = nr;
    // symmetry breaking: constraints placed at the later of their two depths
    if (symmetry && p.n > 1) {
        Symmetry sym;
        query_symmetry(q, MAXAUT, sym);
        if (sym.complete && sym.auts.size() > 1 && sym.less.size() <= MAXSC) {
            std::vector<std::vector<uint32_t>> per(p.n);
            for (auto &c : sym.less) {
                const uint32_t da = p.depth_of[c.first], db = p.depth_of[c.second];
                if (da > db) per[da].push_back(db);            // candidate a < m(b)
                else per[db].push_back(da | 128u);             // candidate b > m(a)
            }
            uint32_t k = 0;
            for (uint32_t d = 0; d < p.n; d++) {
                dp.sc_begin[d] = k;
                for (uint32_t c : per[d]) dp.sc[k++] = c;
            }
            dp.sc_begin[p.n] = k;
            dp.naut = (uint32_t)sym.auts.size();
            for (uint32_t a = 0; a < dp.naut; a++)
                for (uint32_t u = 0; u < p.n; u++) dp.autcol[a][u] = (uint8_t)p.depth_of[sym.auts[a][u]];
        }
    }
    if (p.n == 1 || all_roots) {
        dp.root_base = ALL_VERTICES;
        dp.nroots = h.V;
        return FM_OK;
    }
    // smallest root list among the depth-1 requirements (all anchored at 0)
    uint64_t best = UINT64_MAX;
    uint32_t base = 0;
    uint64_t out_roots = h.out.roots.size();
    for (uint32_t i = dp.req_begin[1]; i < dp.req_begin[1] + (dp.req_begin[2] - dp.req_begin[1] - dp.nself[1]); i++) {
        uint32_t r = dp.req[i];
        uint32_t lid = req_label(r), dir = req_dir(r);
        uint64_t cnt = 0, off = 0;
        if (lid != LABEL_ABSENT) {
            const Csr &c = h.dir(dir);
            cnt = c.root_ptr[lid + 1] - c.root_ptr[lid];
            off = c.root_ptr[lid] + ((dir == 1 && !h.symmetric) ? out_roots : 0);
        }
        if (cnt < best) {
            best = cnt;
            base = (uint32_t)off;
        }
    }
    dp.root_base = base;
    dp.nroots = (uint32_t)(best == UINT64_MAX ? 0 : best);
    return FM_OK;
}

int fm_match_batch(fm_engine *e, const fm_graph *g, const fm_query *const *qs, uint32_t nq,
                   const fm_match_opts *opts, fm_result **out) {
    auto t_start = std::chrono::steady_clock::now();
    if (!e || !g || !out || (nq && !qs)) return fail(FM_ERR_INVALID_ARG, "null argument");
    if (g->device != e->device) return fail(FM_ERR_INVALID_ARG, "graph is not uploaded to the engine's device");
    if (nq == 0 || nq > e->max_queries) return fail(FM_ERR_INVALID_ARG, "query count %u out of range", nq);
    if (e->server) return fail(FM_ERR_BUSY, "the engine is owned by a server");
    fm_match_opts o{};
    if (opts) o = *opts;
    if (o.timeout_ms == 0) o.timeout_ms = 120000;
    if (o.task_rows == 0) o.task_rows = nq == 1 ? FM_QUERY_TASK_ROWS : FM_BATCH_TASK_ROWS;
    if (o.tile_parents > WAVE)
        return fail(FM_ERR_INVALID_ARG, "tile_parents %u is above %u", o.tile_parents, WAVE);
    if (o.tile_parents == 0) o.tile_parents = WAVE;
    if (int rc = check_prealloc(e, g)) return rc;
    const bool default_chunk = o.root_chunk == 0;
    if (o.root_chunk == 0) o.root_chunk = TooLargeThreshold;
    HIP_TRY(hipSetDevice(e->device));
    if (e->dirty) {
        int rc = full_reset(e);
        if (rc) return rc;
    }
    e->epoch++;
    // ---- host-side plans and per-query state ----
    std::vector<DPlan> plans(nq);
    std::vector<DQuery> qh(nq);
    uint64_t chunks = 0;
    // Persistent waves. One query: 8 per CU, of which 4 start at once and the
    // rest join once the launch has run FM_JOIN_MS. Measured on R-MAT 2^20
    // with 16 queries of 6 vertices, one launch each: 77 ms of kernel time
    // per step with 4 waves per CU, 93 ms with 8 from the start (every
    // wave's first bank and page claims and its idle polls contend on shared
    // counters, and a short launch seldom keeps 2,048 waves busy), while the
    // one 30 ms query of the set runs in 21 ms with 8. Several queries:
    // every resident wave (12 per CU on gfx950), all from the start; 1,024
    // queries of 4-12 vertices on the same graph take 340 ms at 12 per CU
    // against 445 ms at 8. An explicit grid_waves starts every wave at once.
    // One query runs the kernel that takes two candidates per lane: on
    // R-MAT 2^20 with 16 queries of 6 vertices, one launch each, 64 ms of
    // kernel time per step against 68 ms.
    // Batches run the same step compiled for 12 waves per CU
    // (fm_match_kernel_batch): at 8 per CU, 1,024 queries took 413 ms.
    const bool paired = nq == 1;
    const uint32_t resident =
        std::max<uint32_t>(1, (uint32_t)(paired ? e->waves_per_cu_paired : e->waves_per_cu_batch) * e->cu_count);
    const uint32_t grid = o.grid_waves ? std::min<uint32_t>(o.grid_waves, resident)
                                       : (nq == 1 ? std::max<uint32_t>(1, std::min(e->grid_waves, resident)) : resident);
    const uint32_t active = (!o.grid_waves && nq == 1 && grid >= 2) ? grid / 2 : grid;
    uint32_t maxn = 1;
    for (uint32_t i = 0; i < nq; i++) {
        if (!qs[i]) return fail(FM_ERR_INVALID_ARG, "null query %u", i);
        int rc = make_dplan(g, qs[i]->h, o.instrument != 0, o.plan, o.symmetry, plans[i]);
        if (rc) return rc;
        if (nq == 1 && default_chunk && !o.waves_per_query) {
            // One query: about 8 root chunks per starting wave, between 64
            // roots (one batch) and TooLargeThreshold, so that waves that
            // draw hub roots do not hold the root depth back. Measured on
            // R-MAT 2^20 with 16 queries of 6 vertices (every query starts
            // from all 2^20 vertices, 1,024 waves): 62.1 ms of kernel time
            // per step with 128-root chunks against 63.6-64.2 ms with 1,024.
            const uint64_t per = ((uint64_t)plans[i].nroots + 8ull * active - 1) / (8ull * active);
            o.root_chunk = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(per, 64), TooLargeThreshold);
        }
        DQuery &Q = qh[i];
        std::memset(&Q, 0, sizeof Q);
        uint32_t n = plans[i].n;
        maxn = std::max(maxn, n);
        Q.page_rows = (uint32_t)(e->page_words / n);
        uint64_t cap_rows = (uint64_t)e->npages * Q.page_rows;
        const bool qtable = o.mode == FM_MODE_TABLE && !qs[i]->count_only;
        Q.count_only = (o.mode == FM_MODE_TABLE && qs[i]->count_only) ? 1u : 0u;
        Q.max_rows = qtable ? (o.max_rows ? std::min<uint64_t>(o.max_rows, cap_rows) : cap_rows) : 0;
        // pages are filled by several waves at once, so a query may hold one
        // partly filled page per wave beyond the full ones
        uint64_t want_pages = (Q.max_rows + Q.page_rows - 1) / Q.page_rows + (o.max_rows ? grid : 0);
        Q.max_pages = qtable ? (uint32_t)std::min<uint64_t>(want_pages, e->npages) : 0;
        Q.max_count = o.max_count;
        Q.max_partials = o.max_partials;
        // roots per claim: root_chunk, or the query's roots cut into
        // waves_per_query claims
        const uint32_t nroots = plans[i].nroots;
        Q.chunk_roots = o.waves_per_query
                            ? std::max<uint32_t>(1, (uint32_t)(((uint64_t)nroots + o.waves_per_query - 1) /
                                                               o.waves_per_query))
                            : o.root_chunk;
        Q.root_chunk_begin = (uint32_t)chunks;
        chunks += (nroots + (uint64_t)Q.chunk_roots - 1) / Q.chunk_roots;
    }
    if (chunks >= (1ull << 32)) return fail(FM_ERR_INVALID_ARG, "too many root chunks");
    // ---- device buffers ----
    if (nq > e->q_cap) {
        (void)hipFree(e->d_plans);
        (void)hipFree(e->d_qs);
        e->d_plans = nullptr;
        e->d_qs = nullptr;
        size_t cap = std::max<size_t>(nq, 64);
        HIP_TRY(hipMalloc((void **)&e->d_plans, cap * sizeof(DPlan)));
        HIP_TRY(hipMalloc((void **)&e->d_qs, cap * sizeof(DQuery)));
        e->q_cap = cap;
    }
    HIP_TRY(hipMemcpyAsync(e->d_plans, plans.data(), nq * sizeof(DPlan), hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync(e->d_qs, qh.data(), nq * sizeof(DQuery), hipMemcpyHostToDevice, e->stream));
    Ctl reset{};
    reset.pending = (long long)chunks;
    reset.root_cursor = active;  // chunks below this are the first claims of the waves that start at once
    reset.done = chunks == 0 ? 1u : 0u;
    reset.timeout_ticks = deadline_ticks(o.timeout_ms, e->wall_hz);
    HIP_TRY(hipMemcpyAsync(e->ctl, &reset, CTL_RESET_BYTES, hipMemcpyHostToDevice, e->stream));
    // ---- launch ----
    KParams P{};
    P.g.V = g->h.V;
    P.g.L = g->h.L();
    P.g.ptr[0] = g->d_ptr[0];
    P.g.ptr[1] = g->d_ptr[1];
    P.g.col = g->d_col[0];
    P.g.skip = g->d_skip;
    P.g.roots = g->d_roots[0];
    P.plans = e->d_plans;
    P.qs = e->d_qs;
    P.ctl = e->ctl;
    P.ring = e->ring;
    P.free_ring = e->free_ring;
    P.bank_ref = e->bank_ref;
    P.banks = e->banks;
    P.pages = e->pages;
    P.page_owner = e->page_owner;
    P.page_fill = e->page_fill;
    P.bank_words = e->bank_words;
    P.page_words = e->page_words;
    P.nq = nq;
    P.total_root_chunks = (uint32_t)chunks;
    P.root_chunk = o.root_chunk;
    P.task_rows = o.task_rows;
    P.tile_parents = o.tile_parents;
    P.table = o.mode == FM_MODE_TABLE ? 1 : 0;
    P.instrument = o.instrument ? 1 : 0;
    P.qcap = e->qcap;
    P.fcap = e->fcap;
    P.nbanks = e->nbanks;
    P.npages = e->npages;
    P.nlevels = std::max<uint32_t>(1, maxn - 1);
    P.starve_ticks = (unsigned long long)(e->wall_hz * 2.0);  // 2 s without a bank
    P.timeout_ticks = reset.timeout_ticks;
    P.active_waves = active < grid ? active : 0;
    P.join_ticks = (unsigned long long)((double)FM_JOIN_MS * (e->wall_hz / 1000.0));
    fm_launch_stats st{};
    st.grid_waves = grid;
    HIP_TRY(hipEventRecord(e->ev0, e->stream));
    if (paired)
        hipLaunchKernelGGL(fm_match_kernel<true>, dim3(grid), dim3(WAVE), 0, e->stream, P);
    else
        hipLaunchKernelGGL(fm_match_kernel_batch, dim3(grid), dim3(WAVE), 0, e->stream, P);
    hipError_t lerr = hipGetLastError();
    HIP_TRY(hipEventRecord(e->ev1, e->stream));
    std::vector<DQuery> back(nq);
    Ctl ctl_back{};
    HIP_TRY(hipMemcpyAsync(back.data(), e->d_qs, nq * sizeof(DQuery), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipMemcpyAsync(&ctl_back, e->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, e->stream));
    hipError_t serr = hipStreamSynchronize(e->stream);
    if (lerr != hipSuccess || serr != hipSuccess) {
        e->dirty = true;
        return fail(FM_ERR_DEVICE, "match kernel failed: %s",
                    hipGetErrorString(lerr != hipSuccess ? lerr : serr));
    }
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, e->ev0, e->ev1));
    auto *r = new fm_result();
    r->e = e;
    r->epoch = e->epoch;
    r->qs = back;
    r->n.resize(nq);
    for (uint32_t i = 0; i < nq; i++) r->n[i] = plans[i].n;
    st.kernel_ms = ms;
    st.tasks_pushed = stat_sum(ctl_back, ST_TASKS);
    st.root_claims = stat_sum(ctl_back, ST_ROOTS);
    st.banks_allocated = stat_sum(ctl_back, ST_BANKS_FRESH);
    st.banks_reused = stat_sum(ctl_back, ST_BANKS_REUSED);
    st.alloc_failures = stat_sum(ctl_back, ST_ALLOC_FAIL);
    st.tasks_split = stat_sum(ctl_back, ST_SPLIT);
    st.requeues = stat_sum(ctl_back, ST_REQUEUE);
    st.pages_allocated = stat_sum(ctl_back, ST_PAGES);
    st.busy_frac = stat_sum(ctl_back, ST_ALIVE) ? (double)stat_sum(ctl_back, ST_BUSY) / (double)stat_sum(ctl_back, ST_ALIVE) : 0.0;
    r->pages_used = std::min(ctl_back.page_bump, e->npages);
    st.status = ctl_back.error;
    if (ctl_back.error) {
        // an aborted launch leaves tasks and bank references behind
        e->dirty = true;
        for (auto &Q : r->qs) {
            if (!Q.status) Q.status = ctl_back.error;
            Q.flags |= FM_RESULT_ABORTED;
        }
    }
    for (auto &Q : r->qs) {
        if (Q.status == FM_ERR_COUNT_LIMIT) Q.flags |= FM_RESULT_ABORTED;
        if (o.mode == FM_MODE_TABLE && !Q.count_only && Q.count > std::min<uint64_t>(Q.rows, Q.max_rows))
            Q.flags |= FM_RESULT_TRUNCATED;
    }
    st.total_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    r->stats = st;
    *out = r;
    if (ctl_back.error == FM_ERR_INTERNAL) {
        const uint32_t *dg = ctl_back.diag;
        return fail(FM_ERR_INTERNAL,
                    "inconsistent task q=%u depth=%u bank=%u begin=%u count=%u skip=%u last_end=%u (n=%u nroots=%u "
                    "bank_words=%llu)",
                    dg[0], dg[1], dg[2], dg[3], dg[4], dg[5], dg[6], dg[7], dg[8], (unsigned long long)e->bank_words);
    }
    if (ctl_back.error == FM_ERR_TIMEOUT)
        return fail(FM_ERR_TIMEOUT, "match launch passed its deadline (%s)", timeout_site(ctl_back).c_str());
    if (ctl_back.error)
        return fail((int)ctl_back.error, "match launch aborted with status %u", ctl_back.error);
    return FM_OK;
}

int fm_match(fm_engine *e, const fm_graph *g, const fm_query *q, const fm_match_opts *opts, fm_result **out) {
    const fm_query *qs[1] = {q};
    if (!q) return fail(FM_ERR_INVALID_ARG, "null query");
    return fm_match_batch(e, g, qs, 1, opts, out);
}

uint32_t fm_result_queries(const fm_result *r) { return r ? (uint32_t)r->qs.size() : 0; }

static uint64_t stored_rows(const DQuery &Q) {
    uint64_t rows = std::min<uint64_t>(Q.rows, Q.max_rows);
    return rows;
}

int fm_result_query(const fm_result *r, uint32_t qi, fm_query_result *out) {
    if (!r || !out || qi >= r->qs.size()) return fail(FM_ERR_INVALID_ARG, "bad result index");
    const DQuery &Q = r->qs[qi];
    std::memset(out, 0, sizeof *out);
    out->count = Q.count;
    out->rows = stored_rows(Q);
    out->status = Q.status;
    out->flags = Q.flags;
    out->vertices = r->n[qi];
    out->alg_bytes = Q.row_words * 4;
    out->row_lookups = Q.lookups;
    for (uint32_t d = 0; d < MAXN; d++) out->partials[d] = Q.partials[d];
    out->candidates = Q.candidates;
    out->probes = Q.probes;
    return FM_OK;
}

int fm_result_stats(const fm_result *r, fm_launch_stats *out) {
    if (!r || !out) return fail(FM_ERR_INVALID_ARG, "null argument");
    *out = r->stats;
    return FM_OK;
}

int fm_result_copy_table(const fm_result *r, uint32_t qi, uint32_t *dst, uint64_t max_rows, int sorted,
                         uint64_t *copied) {
    if (!r || qi >= r->qs.size()) return fail(FM_ERR_INVALID_ARG, "bad result index");
    fm_engine *e = r->e;
    if (r->epoch != e->epoch) return fail(FM_ERR_STALE_RESULT, "the engine has run again; tables are gone");
    const DQuery &Q = r->qs[qi];
    const uint32_t n = r->n[qi];
    uint64_t rows = std::min<uint64_t>(stored_rows(Q), max_rows);
    if (copied) *copied = 0;
    if (rows == 0) return FM_OK;
    if (!dst) return fail(FM_ERR_INVALID_ARG, "null destination");
    HIP_TRY(hipSetDevice(e->device));
    fm_result *rw = const_cast<fm_result *>(r);
    if (rw->page_owner.size() != r->pages_used) {
        rw->page_owner.resize(r->pages_used);
        rw->page_fill.resize(r->pages_used);
        HIP_TRY(hipMemcpy(rw->page_owner.data(), e->page_owner, r->pages_used * 4ull, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(rw->page_fill.data(), e->page_fill, r->pages_used * 4ull, hipMemcpyDeviceToHost));
    }
    uint64_t done = 0;
    for (uint32_t p = 0; p < r->pages_used && done < rows; p++) {
        if (r->page_owner[p] != qi || r->page_fill[p] == 0) continue;
        uint64_t take = std::min<uint64_t>(r->page_fill[p], rows - done);
        HIP_TRY(hipMemcpy(dst + done * n, e->pages + (size_t)p * e->page_words, take * n * 4, hipMemcpyDeviceToHost));
        done += take;
    }
    if (sorted && done > 1) {
        std::vector<uint64_t> idx(done);
        std::iota(idx.begin(), idx.end(), 0);
        std::sort(idx.begin(), idx.end(), [&](uint64_t a, uint64_t b) {
            const uint32_t *x = dst + a * n, *y = dst + b * n;
            for (uint32_t c = 0; c < n; c++)
                if (x[c] != y[c]) return x[c] < y[c];
            return false;
        });
        std::vector<uint32_t> tmp(done * n);
        for (uint64_t i = 0; i < done; i++) std::memcpy(&tmp[i * n], dst + idx[i] * n, n * 4);
        std::memcpy(dst, tmp.data(), done * n * 4);
    }
    if (copied) *copied = done;
    return FM_OK;
}

void fm_result_destroy(fm_result *r) { delete r; }

}  // extern "C"

// ---------------------------------------------------------------------------
// Serving host side.

struct fm_server {
    fm_engine *e = nullptr;
    const fm_graph *g = nullptr;
    fm_server_config cfg{};
    SrvShared *sh = nullptr;      // pinned, coherent
    DPlan *hplans = nullptr;      // pinned
    SrvQInit *hqinit = nullptr;   // pinned
    SrvDone *hdone = nullptr;     // pinned
    uint32_t *slot_pages = nullptr;
    hipStream_t copy_stream = nullptr;
    hipEvent_t kernel_done = nullptr;
    bool launched = false;
    int failed = 0;  // status of the launch that stopped the server (0: running)
    std::chrono::steady_clock::time_point t0;
    std::vector<uint32_t> free_slots;
    std::vector<uint32_t> slot_ticket;   // ticket in the slot (0 = none)
    std::vector<uint8_t> slot_state;     // 0 free, 1 in flight, 2 done
    std::vector<double> slot_submit_ms;
    std::vector<uint32_t> slot_n;
    uint32_t next_ticket = 1;
    uint32_t sub_tail = 0;
    fm_server_stats st{};
};

static double server_ms(const fm_server *s) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - s->t0).count();
}

static int server_launch(fm_server *s) {
    fm_engine *e = s->e;
    HIP_TRY(hipSetDevice(e->device));
    if (e->dirty) {
        int rc = full_reset(e);
        if (rc) return rc;
    }
    Ctl reset{};
    reset.timeout_ticks = (unsigned long long)(s->cfg.max_launch_ms * (e->wall_hz / 1000.0));
    HIP_TRY(hipMemcpyAsync(e->ctl, &reset, CTL_RESET_BYTES, hipMemcpyHostToDevice, e->stream));
    // the previous launch (finished: its event was reached) left the state
    // word closed; reopen it for this one
    unsigned long long st = 0;
    HIP_TRY(hipMemcpyAsync(&st, &e->ctl->srv_state, sizeof st, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (st & SRV_CLOSING) {
        st &= ~SRV_CLOSING;
        HIP_TRY(hipMemcpyAsync(&e->ctl->srv_state, &st, sizeof st, hipMemcpyHostToDevice, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
    }
    KParams P{};
    const fm_graph *g = s->g;
    P.g.V = g->h.V;
    P.g.L = g->h.L();
    P.g.ptr[0] = g->d_ptr[0];
    P.g.ptr[1] = g->d_ptr[1];
    P.g.col = g->d_col[0];
    P.g.skip = g->d_skip;
    P.g.roots = g->d_roots[0];
    P.plans = e->d_plans;
    P.plans_rw = e->d_plans;
    P.qs = e->d_qs;
    P.ctl = e->ctl;
    P.ring = e->ring;
    P.free_ring = e->free_ring;
    P.bank_ref = e->bank_ref;
    P.banks = e->banks;
    P.pages = e->pages;
    P.page_owner = e->page_owner;
    P.page_fill = e->page_fill;
    P.bank_words = e->bank_words;
    P.page_words = e->page_words;
    P.nq = s->cfg.slots;
    P.total_root_chunks = 0;
    P.root_chunk = s->cfg.root_chunk;
    P.task_rows = s->cfg.task_rows;
    P.tile_parents = WAVE;
    P.table = s->cfg.table ? 1 : 0;
    P.qcap = e->qcap;
    P.fcap = e->fcap;
    P.nbanks = e->nbanks;
    P.npages = e->npages;
    P.nlevels = NLEVELS - 1;
    P.starve_ticks = (unsigned long long)(e->wall_hz * 2.0);
    P.timeout_ticks = reset.timeout_ticks;
    P.serving = 1;
    P.srv = s->sh;
    P.srv_plans = s->hplans;
    P.srv_qinit = s->hqinit;
    P.srv_done = s->hdone;
    P.slot_pages = s->slot_pages;
    P.pfree_ring = e->pfree_ring;
    P.pfcap = e->pfcap;
    P.max_slots = s->cfg.slots;
    P.max_slot_pages = s->cfg.slot_pages;
    P.idle_ticks = (unsigned long long)(s->cfg.idle_ms * (e->wall_hz / 1000.0));
    P.slice_ticks = (unsigned long long)(s->cfg.slice_ms * (e->wall_hz / 1000.0));
    hipLaunchKernelGGL(fm_match_kernel<false>, dim3(e->grid_waves), dim3(WAVE), 0, e->stream, P);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(s->kernel_done, e->stream));
    s->launched = true;
    s->st.launches++;
    return FM_OK;
}

// A request (admitted or not) whose completion the device has not
// published.
static bool server_has_work(const fm_server *s) {
    for (size_t i = 0; i < s->slot_state.size(); i++)
        if (s->slot_state[i] == 1 && __atomic_load_n(&s->hdone[i].ticket, __ATOMIC_ACQUIRE) != s->slot_ticket[i])
            return true;
    return false;
}

// Relaunches the persistent kernel when it has gone idle or retired and
// work waits.
static int server_pump(fm_server *s) {
    // After an aborted launch the device queues, pools and submission head
    // are no longer trustworthy: the server stays stopped.
    if (s->failed) return fail(s->failed, "the server stopped after an earlier error (status %d)", s->failed);
    if (s->launched) {
        hipError_t q = hipEventQuery(s->kernel_done);
        if (q == hipErrorNotReady) return FM_OK;
        if (q != hipSuccess) {
            s->launched = false;
            s->failed = FM_ERR_DEVICE;
            return fail(FM_ERR_DEVICE, "serving kernel failed: %s", hipGetErrorString(q));
        }
        s->launched = false;
        Ctl back{};
        HIP_TRY(hipMemcpy(&back, s->e->ctl, CTL_RESET_BYTES, hipMemcpyDeviceToHost));
        s->st.dropped_tasks += stat_sum(back, ST_STALE);
        if (back.error) {
            s->e->dirty = true;
            s->failed = (int)back.error;
            if (back.error == FM_ERR_TIMEOUT)
                return fail(FM_ERR_TIMEOUT, "serving kernel passed its deadline (%s)", timeout_site(back).c_str());
            if (back.error == FM_ERR_INTERNAL)
                return fail(FM_ERR_INTERNAL,
                            "serving kernel caught an inconsistent task q=%u tag=%u depth=%u bank=%u begin=%u "
                            "count=%u skip=%u last_end=%u (n=%u nroots=%u; slot ticket %u, pending %d)",
                            back.diag[0] & 0xFFFFFF, back.diag[0] >> 24, back.diag[1], back.diag[2], back.diag[3],
                            back.diag[4], back.diag[5], back.diag[6], back.diag[7], back.diag[8], back.diag[9],
                            (int)back.diag[10]);
            return fail((int)back.error, "serving kernel aborted with status %u", back.error);
        }
    }
    if (server_has_work(s)) return server_launch(s);
    return FM_OK;
}

extern "C" {

int fm_server_create(fm_engine *e, const fm_graph *g, const fm_server_config *cfg, fm_server **out) {
    if (!e || !g || !out) return fail(FM_ERR_INVALID_ARG, "null argument");
    if (e->server) return fail(FM_ERR_BUSY, "the engine already has a server");
    if (g->device != e->device) return fail(FM_ERR_INVALID_ARG, "graph is not uploaded to the engine's device");
    if (int rc = check_prealloc(e, g)) return rc;
    auto *s = new fm_server();
    s->e = e;
    s->g = g;
    if (cfg) s->cfg = *cfg;
    if (!s->cfg.slots) s->cfg.slots = 256;
    if (!s->cfg.slot_pages) s->cfg.slot_pages = 256;
    if (!s->cfg.idle_ms) s->cfg.idle_ms = 20;
    if (!s->cfg.root_chunk) s->cfg.root_chunk = TooLargeThreshold;
    if (!s->cfg.task_rows) s->cfg.task_rows = SubMaxMulSpill;
    if (!s->cfg.bank_budget) s->cfg.bank_budget = e->nbanks;
    if (!s->cfg.max_launch_ms) s->cfg.max_launch_ms = 120000;
    if (!s->cfg.slice_ms) s->cfg.slice_ms = 2000;
    if (s->cfg.slice_ms >= s->cfg.max_launch_ms)
        return fail(FM_ERR_INVALID_ARG, "slice_ms %u must be below max_launch_ms %u", s->cfg.slice_ms,
                    s->cfg.max_launch_ms);
    if (!cfg) s->cfg.table = 1;
    const uint32_t S = s->cfg.slots;
    auto fail_out = [&](int rc) {
        s->e = nullptr;  // not owned yet
        delete s;
        return rc;
    };
    if (S > e->max_queries) return fail_out(fail(FM_ERR_INVALID_ARG, "slots %u exceed the engine's max_queries", S));
    if (S > SUB_CAP) return fail_out(fail(FM_ERR_INVALID_ARG, "slots %u exceed the submission ring (%u)", S, SUB_CAP));
    HIP_TRY(hipSetDevice(e->device));
    if (hipHostMalloc((void **)&s->sh, sizeof(SrvShared), hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess ||
        hipHostMalloc((void **)&s->hplans, S * sizeof(DPlan), hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess ||
        hipHostMalloc((void **)&s->hqinit, S * sizeof(SrvQInit), hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess ||
        hipHostMalloc((void **)&s->hdone, S * sizeof(SrvDone), hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess)
        return fail_out(fail(FM_ERR_OUT_OF_MEMORY, "pinned serving buffers"));
    std::memset(s->sh, 0, sizeof(SrvShared));
    std::memset(s->hplans, 0, S * sizeof(DPlan));
    std::memset(s->hqinit, 0, S * sizeof(SrvQInit));
    std::memset(s->hdone, 0, S * sizeof(SrvDone));
    if (hipMalloc((void **)&s->slot_pages, (size_t)S * s->cfg.slot_pages * 4) != hipSuccess ||
        hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&s->kernel_done, hipEventDisableTiming) != hipSuccess)
        return fail_out(fail(FM_ERR_OUT_OF_MEMORY, "serving device buffers"));
    // slot state on the device starts zeroed (no pages to recycle); page
    // lists start empty
    HIP_TRY(hipMemset(s->slot_pages, 0xFF, (size_t)S * s->cfg.slot_pages * 4));
    if (S > e->q_cap) {
        (void)hipFree(e->d_plans);
        (void)hipFree(e->d_qs);
        e->d_plans = nullptr;
        e->d_qs = nullptr;
        HIP_TRY(hipMalloc((void **)&e->d_plans, S * sizeof(DPlan)));
        HIP_TRY(hipMalloc((void **)&e->d_qs, S * sizeof(DQuery)));
        e->q_cap = S;
    }
    HIP_TRY(hipMemset(e->d_qs, 0, S * sizeof(DQuery)));
    if (!e->pfree_ring) {
        e->pfcap = pow2_at_least(e->npages);
        HIP_TRY(hipMalloc((void **)&e->pfree_ring, (size_t)e->pfcap * sizeof(FSlot)));
    }
    int rc = full_reset(e);  // also clears the serving counters
    if (rc) return fail_out(rc);
    s->slot_ticket.assign(S, 0);
    s->slot_state.assign(S, 0);
    s->slot_submit_ms.assign(S, 0);
    s->slot_n.assign(S, 0);
    for (uint32_t i = 0; i < S; i++) s->free_slots.push_back(S - 1 - i);
    s->t0 = std::chrono::steady_clock::now();
    e->server = s;
    *out = s;
    return FM_OK;
}

static int slot_of(const fm_server *s, uint32_t ticket) {
    for (size_t i = 0; i < s->slot_ticket.size(); i++)
        if (s->slot_ticket[i] == ticket && s->slot_state[i]) return (int)i;
    return -1;
}

int fm_server_submit(fm_server *s, const fm_query *q, uint64_t max_rows, uint32_t *ticket) {
    if (!s || !q || !ticket) return fail(FM_ERR_INVALID_ARG, "null argument");
    // The device publishes live_banks at admissions and completions; with
    // nothing in flight the budget cannot be held by this server's requests,
    // so an idle server always admits. Every submission holds a slot until
    // it completes and slots <= SUB_CAP, so the ring cannot overrun.
    uint32_t live = __atomic_load_n(&s->sh->live_banks, __ATOMIC_RELAXED);
    const bool idle = s->free_slots.size() == s->slot_state.size();
    if (s->free_slots.empty() || (!idle && live >= s->cfg.bank_budget)) {
        s->st.refused++;
        if (!s->free_slots.empty()) s->st.refused_memory++;
        return fail(FM_ERR_BUSY, "no slot or bank budget available");
    }
    uint32_t slot = s->free_slots.back();
    DPlan dp;
    int rc = make_dplan(s->g, q->h, false, s->cfg.plan, s->cfg.symmetry, dp);
    if (rc) return rc;
    const uint32_t n = dp.n;
    SrvQInit qi{};
    qi.page_rows = (uint32_t)(s->e->page_words / n);
    qi.max_pages = s->cfg.table ? s->cfg.slot_pages : 0;
    uint64_t cap = (uint64_t)qi.max_pages * qi.page_rows;
    qi.max_rows = s->cfg.table ? (max_rows ? std::min<uint64_t>(max_rows, cap) : cap) : 0;
    uint32_t t = s->next_ticket++;
    if (s->next_ticket == 0) s->next_ticket = 1;
    qi.ticket = t;
    s->free_slots.pop_back();
    std::memcpy(&s->hplans[slot], &dp, sizeof dp);
    std::memcpy(&s->hqinit[slot], &qi, sizeof qi);
    __atomic_store_n(&s->hdone[slot].ticket, 0u, __ATOMIC_RELAXED);
    SrvSub sub{slot, t, {0, 0}};
    std::memcpy(&s->sh->sub[s->sub_tail % SUB_CAP], &sub, sizeof sub);
    s->sub_tail++;
    __atomic_store_n(&s->sh->sub_tail, s->sub_tail, __ATOMIC_RELEASE);
    s->slot_ticket[slot] = t;
    s->slot_state[slot] = 1;
    s->slot_submit_ms[slot] = server_ms(s);
    s->slot_n[slot] = n;
    s->st.submitted++;
    *ticket = t;
    return server_pump(s);
}

int fm_server_poll(fm_server *s, fm_server_completion *out, uint32_t max, uint32_t *n) {
    if (!s || !n) return fail(FM_ERR_INVALID_ARG, "null argument");
    *n = 0;
    for (size_t i = 0; i < s->slot_state.size() && *n < max; i++) {
        if (s->slot_state[i] != 1) continue;
        SrvDone *d = &s->hdone[i];
        if (__atomic_load_n(&d->ticket, __ATOMIC_ACQUIRE) != s->slot_ticket[i]) continue;
        fm_server_completion c{};
        c.ticket = s->slot_ticket[i];
        c.status = __atomic_load_n(&d->status, __ATOMIC_RELAXED);
        c.flags = __atomic_load_n(&d->flags, __ATOMIC_RELAXED);
        c.vertices = s->slot_n[i];
        c.count = ((uint64_t)__atomic_load_n(&d->count_hi, __ATOMIC_RELAXED) << 32) | d->count_lo;
        c.rows = ((uint64_t)__atomic_load_n(&d->rows_hi, __ATOMIC_RELAXED) << 32) | d->rows_lo;
        if (s->cfg.table && c.count > c.rows) c.flags |= FM_RESULT_TRUNCATED;
        c.submit_ms = s->slot_submit_ms[i];
        c.done_ms = server_ms(s);
        s->slot_state[i] = 2;
        s->st.completed++;
        if (out) out[*n] = c;
        (*n)++;
    }
    return server_pump(s);
}

int fm_server_copy_table(fm_server *s, uint32_t ticket, uint32_t *dst, uint64_t max_rows, int sorted,
                         uint64_t *copied) {
    if (!s) return fail(FM_ERR_INVALID_ARG, "null server");
    int slot = slot_of(s, ticket);
    if (slot < 0 || s->slot_state[slot] != 2) return fail(FM_ERR_INVALID_ARG, "ticket %u is not complete", ticket);
    if (copied) *copied = 0;
    fm_engine *e = s->e;
    const SrvDone *d = &s->hdone[slot];
    const uint32_t n = s->slot_n[slot];
    uint64_t rows = std::min<uint64_t>(((uint64_t)d->rows_hi << 32) | d->rows_lo, max_rows);
    uint32_t npg = d->npages;
    if (rows == 0 || npg == 0) return FM_OK;
    if (!dst) return fail(FM_ERR_INVALID_ARG, "null destination");
    HIP_TRY(hipSetDevice(e->device));
    // dense table: row r in slot page r / page_rows
    const uint64_t pr = s->hqinit[slot].page_rows;
    std::vector<uint32_t> ids(npg);
    HIP_TRY(hipMemcpyAsync(ids.data(), s->slot_pages + (size_t)slot * s->cfg.slot_pages, npg * 4ull,
                           hipMemcpyDeviceToHost, s->copy_stream));
    HIP_TRY(hipStreamSynchronize(s->copy_stream));
    uint64_t done = 0;
    for (uint32_t k = 0; k < npg && done < rows; k++) {
        const uint32_t p = ids[k];
        if (p >= e->npages) return fail(FM_ERR_INTERNAL, "ticket %u: table page %u is missing", ticket, k);
        const uint64_t take = std::min<uint64_t>(pr, rows - done);
        HIP_TRY(hipMemcpyAsync(dst + done * n, e->pages + (size_t)p * e->page_words, take * n * 4ull,
                               hipMemcpyDeviceToHost, s->copy_stream));
        done += take;
    }
    HIP_TRY(hipStreamSynchronize(s->copy_stream));
    if (sorted && done > 1) {
        std::vector<uint64_t> idx(done);
        std::iota(idx.begin(), idx.end(), 0);
        std::sort(idx.begin(), idx.end(), [&](uint64_t a, uint64_t b) {
            const uint32_t *x = dst + a * n, *y = dst + b * n;
            for (uint32_t c = 0; c < n; c++)
                if (x[c] != y[c]) return x[c] < y[c];
            return false;
        });
        std::vector<uint32_t> tmp(done * n);
        for (uint64_t i = 0; i < done; i++) std::memcpy(&tmp[i * n], dst + idx[i] * n, n * 4);
        std::memcpy(dst, tmp.data(), done * n * 4);
    }
    if (copied) *copied = done;
    return FM_OK;
}

int fm_server_release(fm_server *s, uint32_t ticket) {
    if (!s) return fail(FM_ERR_INVALID_ARG, "null server");
    int slot = slot_of(s, ticket);
    if (slot < 0 || s->slot_state[slot] != 2) return fail(FM_ERR_INVALID_ARG, "ticket %u is not complete", ticket);
    s->slot_state[slot] = 0;
    s->slot_ticket[slot] = 0;
    s->free_slots.push_back((uint32_t)slot);
    return FM_OK;
}

int fm_server_stats_get(fm_server *s, fm_server_stats *out) {
    if (!s || !out) return fail(FM_ERR_INVALID_ARG, "null argument");
    *out = s->st;
    uint32_t inflight = 0;
    for (auto st : s->slot_state) inflight += st == 1;
    out->in_flight = inflight;
    out->live_banks = __atomic_load_n(&s->sh->live_banks, __ATOMIC_RELAXED);
    return FM_OK;
}

void fm_server_destroy(fm_server *s) {
    if (!s) return;
    if (s->sh) __atomic_store_n(&s->sh->shutdown, 1u, __ATOMIC_RELEASE);
    if (s->sh && s->e && !s->failed) {
        // give in-flight queries 10 s to drain (relaunching retired
        // launches), then stop the launch
        auto t_stop = std::chrono::steady_clock::now() + std::chrono::seconds(10);
        while (std::chrono::steady_clock::now() < t_stop) {
            if (s->launched && hipEventQuery(s->kernel_done) == hipErrorNotReady) {
                std::this_thread::sleep_for(std::chrono::milliseconds(1));
                continue;
            }
            if (server_pump(s) != FM_OK || !s->launched) break;
        }
        if (s->launched) {
            if (hipEventQuery(s->kernel_done) == hipErrorNotReady)
                __atomic_store_n(&s->sh->abort, 1u, __ATOMIC_RELEASE);
            (void)hipEventSynchronize(s->kernel_done);
        }
    }
    if (s->e) {
        (void)hipSetDevice(s->e->device);
        s->e->server = nullptr;
        s->e->dirty = true;  // batch launches re-initialise the shared pools
    }
    if (s->copy_stream) (void)hipStreamDestroy(s->copy_stream);
    if (s->kernel_done) (void)hipEventDestroy(s->kernel_done);
    (void)hipFree(s->slot_pages);
    (void)hipHostFree(s->sh);
    (void)hipHostFree(s->hplans);
    (void)hipHostFree(s->hqinit);
    (void)hipHostFree(s->hdone);
    delete s;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Intersection host side.

// The no-allocation entry point's work array (claim counter, pairs' first
// segments, per-segment status words): a fixed device array, each GPU
// holding its own copy until the process ends. It starts zeroed (a static
// device array), which the status words' call epochs need. Host code uses
// the address hipGetSymbolAddress returns for the device, taken once per
// device at engine creation, never the array's name.
constexpr uint32_t NOALLOC_WORDS = FM_NOALLOC_SCRATCH_BYTES / 4;
// its fixed geometry: room for the largest pair count's table, the rest status words
constexpr uint64_t NOALLOC_PAIRS = (uint64_t)FM_NOALLOC_MAX_PAIRS + 1;
constexpr uint32_t NOALLOC_SEGS = (uint32_t)((NOALLOC_WORDS - ISX_HDR - NOALLOC_PAIRS) / 2);
static_assert(isx_work_words(NOALLOC_PAIRS, NOALLOC_SEGS) <= NOALLOC_WORDS, "the work array fits the scratch");
static __device__ __align__(256) uint32_t scratch[NOALLOC_WORDS];
constexpr int FM_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static uint32_t *g_scratch_ptr[FM_MAX_DEVICES];

// Called by fm_engine_create with the engine's device current.
static int isect_scratch_init(int device) {
    if (device < 0 || device >= FM_MAX_DEVICES) return fail(FM_ERR_INVALID_ARG, "device %d out of range", device);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return FM_OK;
    void *ptr = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (rc != hipSuccess || !ptr)
        return fail(FM_ERR_DEVICE, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(rc));
    g_scratch_ptr[device] = (uint32_t *)ptr;
    return FM_OK;
}

static uint32_t *isect_scratch_of(int device) {
    if (device < 0 || device >= FM_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[device];
}

// The plan and main kernels on the engine's stream with `work` holding
// isx_work_words(npairs + 1, cap) words laid out for cap segments;
// allocates nothing. When the pairs need more than cap segments the kernels
// do no work and *over is set.
static int isect_run(fm_engine *e, uint32_t *work, uint32_t cap, const uint32_t *data, const uint64_t *a_off,
                     const uint32_t *a_len, const uint64_t *b_off, const uint32_t *b_len, uint32_t npairs,
                     uint32_t *counts, const uint64_t *out_off, uint32_t *out, double *kernel_ms, bool *over) {
    if (!data || !a_off || !a_len || !b_off || !b_len || !counts) return fail(FM_ERR_INVALID_ARG, "null argument");
    if (out_off && !out) return fail(FM_ERR_INVALID_ARG, "out_off without out");
    if (kernel_ms) *kernel_ms = 0;
    if (over) *over = false;
    if (npairs == 0) return FM_OK;
    HIP_TRY(hipSetDevice(e->device));
    HIP_TRY(hipEventRecord(e->ev0, e->stream));
    isx_launch(e->stream, (uint32_t)e->cu_count, work, cap, data, a_off, a_len, b_off, b_len, npairs, counts, out_off, out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(e->ev1, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, e->ev0, e->ev1));
    if (kernel_ms) *kernel_ms = ms;
    uint32_t flags = 0;
    HIP_TRY(hipMemcpy(&flags, work + ISX_FLAGS, 4, hipMemcpyDeviceToHost));
    if (flags & ISX_GAVE_UP)
        return fail(FM_ERR_INTERNAL, "intersection: a segment's wait for the earlier segments of its pair passed its limit");
    if (over) *over = (flags & ISX_OVER) != 0;
    return FM_OK;
}

extern "C" {

int fm_intersect_pairs(fm_engine *e, const uint32_t *data, const uint64_t *a_off, const uint32_t *a_len,
                       const uint64_t *b_off, const uint32_t *b_len, uint32_t npairs, uint32_t *counts,
                       const uint64_t *out_off, uint32_t *out, double *kernel_ms) {
    if (!e) return fail(FM_ERR_INVALID_ARG, "null argument");
    if (npairs == 0 || !a_len || !b_len)
        return isect_run(e, nullptr, 0, data, a_off, a_len, b_off, b_len, npairs, counts, out_off, out, kernel_ms,
                         nullptr);
    // size the work array: the segments of every pair (lengths read back)
    HIP_TRY(hipSetDevice(e->device));
    std::vector<uint32_t> la(npairs), lb(npairs);
    HIP_TRY(hipMemcpy(la.data(), a_len, npairs * 4ull, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(lb.data(), b_len, npairs * 4ull, hipMemcpyDeviceToHost));
    uint64_t segs = 0;
    for (uint32_t p = 0; p < npairs; p++) segs += isx_segments(la[p], lb[p]);
    if (segs >= (1ull << 31)) return fail(FM_ERR_INVALID_ARG, "too many intersection segments");
    if (segs > e->isect_segs || npairs + 1ull > e->isect_pairs) {
        // a new array, laid out for at least as many segments and pairs as
        // the old one (one layout for the array's life)
        const uint64_t nsegs = std::max<uint64_t>(segs, e->isect_segs), npr = std::max<uint64_t>(npairs + 1ull, e->isect_pairs);
        const size_t bytes = isx_work_words(npr, nsegs) * 4;
        (void)hipFree(e->isect_buf);
        e->isect_buf = nullptr;
        e->isect_segs = e->isect_pairs = 0;
        HIP_TRY(hipMalloc((void **)&e->isect_buf, bytes));
        // status words carry a call epoch: a new array starts from zero
        HIP_TRY(hipMemset(e->isect_buf, 0, bytes));
        e->isect_segs = nsegs, e->isect_pairs = npr;
    }
    return isect_run(e, (uint32_t *)e->isect_buf, (uint32_t)e->isect_segs, data, a_off, a_len, b_off, b_len, npairs,
                     counts, out_off, out, kernel_ms, nullptr);
}

int fm_intersect_pairs_noalloc(fm_engine *e, const uint32_t *data, const uint64_t *a_off, const uint32_t *a_len,
                               const uint64_t *b_off, const uint32_t *b_len, uint32_t npairs, uint32_t *counts,
                               const uint64_t *out_off, uint32_t *out, double *kernel_ms) {
    if (!e) return fail(FM_ERR_INVALID_ARG, "null argument");
    if (npairs > FM_NOALLOC_MAX_PAIRS)
        return fail(FM_ERR_INVALID_ARG, "%u pairs exceed the no-allocation limit of %u pairs", npairs,
                    (uint32_t)FM_NOALLOC_MAX_PAIRS);
    uint32_t *work = isect_scratch_of(e->device);
    if (!work) return fail(FM_ERR_DEVICE, "no scratch address for device %d", e->device);
    const uint32_t cap = NOALLOC_SEGS;
    bool over = false;
    int rc = isect_run(e, work, cap, data, a_off, a_len, b_off, b_len, npairs, counts, out_off, out, kernel_ms, &over);
    if (rc) return rc;
    if (over)
        return fail(FM_ERR_INVALID_ARG,
                    "the pairs need more than the %u segments of the no-allocation scratch (%u bytes)", cap,
                    (uint32_t)FM_NOALLOC_SCRATCH_BYTES);
    return FM_OK;
}

int fm_device_alloc(uint64_t bytes, void **out) {
    if (!out) return fail(FM_ERR_INVALID_ARG, "null argument");
    HIP_TRY(hipMalloc(out, std::max<uint64_t>(bytes, 4)));
    return FM_OK;
}

int fm_device_free(void *p) {
    HIP_TRY(hipFree(p));
    return FM_OK;
}

int fm_copy_to_device(void *dst, const void *src, uint64_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return FM_OK;
}

int fm_copy_to_host(void *dst, const void *src, uint64_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return FM_OK;
}

}  // extern "C"
