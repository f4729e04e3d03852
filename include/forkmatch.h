/*
 * ForkMatch for MI355X: the C ABI of libforkmatch.so.
 *
 * The reference exposes no library API: its surface is the CLI
 * `ForkMatch <query.g> <data.g>` (src/smsm.cu:1351-1385), the `.g` grammar
 * (README.md:7-25) and the knobs of src/environment.cuh. Its internal
 * host->device boundary is one kernel launch,
 *   initPosting<256,sym><<<G,256>>>(MemoryPool, FuncPair* forest,
 *       GPUGraph* data, Requirements reqs, unsigned* output)
 * (src/smsm.cu:1305-1326), preceded by the host loaders parseGraph
 * (src/trie.cuh:1051-1222), fileParse (src/hostparser.cuh:286-509) and
 * preProcessQuery (src/ccsr.cuh:485-728). Each entry point below names the
 * reference function it replaces. Plain pointers and sizes only: no C++ or
 * torch types cross this boundary.
 *
 * Error convention: every function returns an int status (FM_OK = 0). The
 * reference calls exit(code); the numeric codes of its parse errors are kept
 * so the CLI can exit with the same values. fm_last_error() returns a text
 * description of the last failure on the calling thread.
 *
 * Threading: one host thread per fm_engine. Graphs are immutable after
 * construction and may be shared by several engines.
 */
#ifndef FORKMATCH_H
#define FORKMATCH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
enum fm_status {
    FM_OK = 0,
    /* Disconnected query, plan check failure, arena overflow
     * (src/ccsr.cuh:258, src/ccsr.cuh:719-725, src/trie.cuh:1044). */
    FM_ERR_GENERIC = 1,
    /* Data graph file missing (src/trie.cuh:1057-1062). */
    FM_ERR_DATA_MISSING_FILE = 12,
    /* Data graph does not start with "t # 0\n" (src/trie.cuh:1075-1078). */
    FM_ERR_DATA_BAD_START = 76,
    /* Data graph does not end with "t # -1" (src/trie.cuh:1080-1083). */
    FM_ERR_DATA_BAD_END = 77,
    /* Malformed data graph header (src/trie.cuh:1088-1092). */
    FM_ERR_DATA_BAD_HEADER = 78,
    /* Malformed data graph body line or vertex id out of range. The
     * reference traps on the device here (src/trie.cuh:845-856). */
    FM_ERR_DATA_BAD_BODY = 79,
    /* Wrong argument count on the CLI (src/smsm.cu:1378-1385). */
    FM_ERR_ARGV = 99,
    /* Query file missing / invalid / short (src/hostparser.cuh:15-17). */
    FM_ERR_QUERY_MISSING_FILE = 232,
    FM_ERR_QUERY_INVALID = 233,
    FM_ERR_QUERY_MISSING_DATA = 234,
    /* Statuses with no reference counterpart (the reference __trap()s). */
    FM_ERR_INVALID_ARG = 300,
    FM_ERR_DEVICE = 301,          /* HIP runtime error */
    FM_ERR_OUT_OF_MEMORY = 302,   /* hipMalloc failed */
    FM_ERR_POOL_EXHAUSTED = 303,  /* no bank could be allocated for too long */
    FM_ERR_QUEUE_FULL = 304,      /* a device work queue overflowed */
    FM_ERR_TIMEOUT = 305,         /* the launch passed its deadline */
    FM_ERR_COUNT_LIMIT = 306,     /* a query passed opts.max_count */
    FM_ERR_QUERY_TOO_LARGE = 307, /* query beyond the device plan limits */
    FM_ERR_STALE_RESULT = 308,    /* the engine ran again; tables are gone */
    FM_ERR_NOT_SUPPORTED = 309,
    FM_ERR_INTERNAL = 310         /* the device caught an inconsistent task */
};

/* Query size limit of the device plan (environment.h FM_MAX_QUERY_VERTICES). */
#define FM_MAX_QUERY_VERTICES_ABI 32

/* Per-query result flags (fm_query_result.flags). */
#define FM_RESULT_TRUNCATED 1u   /* table capacity reached; count is exact */
#define FM_RESULT_ABORTED 2u     /* the launch aborted; count is partial */

/* ---- opaque handles --------------------------------------------------- */
typedef struct fm_graph fm_graph;   /* device-resident data graph */
typedef struct fm_query fm_query;   /* parsed query + matching plan */
typedef struct fm_engine fm_engine; /* device context: pools, queues, stream */
typedef struct fm_result fm_result; /* results of one match launch */

/* ---- library ----------------------------------------------------------- */
const char *fm_last_error(void);
const char *fm_version(void);
/* Number of visible HIP devices (0 without a GPU; never fails). */
int fm_device_count(void);

/* ---- data graph ---------------------------------------------------------
 * Replaces parseGraph + parseCSR + csrToTrie (src/trie.cuh:800-951,
 * 1051-1222): the data graph becomes a label-partitioned CSR in HBM.
 * Duplicate edge lines are collapsed; vertex labels are read and ignored
 * (the reference stores them write-only, src/trie.cuh:872). */
typedef struct fm_graph_info {
    uint32_t vertices;      /* V */
    uint32_t labels;        /* distinct edge labels L */
    uint64_t edges;         /* directed edges after de-duplication */
    uint32_t symmetric;     /* 1 if every (u,v,l) has its (v,u,l) */
    uint32_t header_edges;  /* edge count stated in the .g header */
    uint64_t device_bytes;  /* HBM held by the graph */
} fm_graph_info;

/* Host-side graph (parse + CSR build only, no device). Used by tools that
 * convert or inspect graphs, and by the tests on machines without a GPU. */
int fm_graph_parse_g(const char *path, fm_graph **out);
int fm_graph_from_edges(uint32_t vertices, uint64_t nedges,
                        const uint32_t *src, const uint32_t *dst,
                        const uint32_t *label, fm_graph **out);
/* Uploads the CSR to the current HIP device (idempotent). */
int fm_graph_upload(fm_graph *g);
/* Load a .g data file onto the current HIP device (replaces parseGraph,
 * trie.cuh:1051-1222, and parseCSR, trie.cuh:800-951). The host checks the
 * frame and header; the body's lines are found and parsed on the device
 * (one thread per line), with the same checks, error codes and messages as
 * fm_graph_parse_g and 64-bit offsets throughout (no 2 GiB limit); the CSR
 * is built (sorted, de-duplicated, root lists and row statistics) and
 * checked for symmetry on the device, and copied to the host. A directed
 * graph, or one with a label value of 2^24 or more, is built on the host. */
int fm_graph_load_g(const char *path, fm_graph **out);
int fm_graph_info_get(const fm_graph *g, fm_graph_info *out);
/* Copies the host CSR. ptr has labels*vertices+1 entries (row (l,v) is
 * col[ptr[l*V+v] .. ptr[l*V+v+1])); col has edges entries. label_values
 * maps dense label id -> label value as written in the file. Any pointer
 * may be NULL. dir 0 = out-rows, 1 = in-rows. */
int fm_graph_copy_csr(const fm_graph *g, int dir, uint32_t *ptr,
                      uint32_t *col, uint32_t *label_values);
/* Copies the device CSR back (parity check of the upload). */
int fm_graph_copy_device_csr(const fm_graph *g, int dir, uint32_t *ptr,
                             uint32_t *col);
/* Binary graph cache: the .g text of the large configurations exceeds the
 * reference's 2 GiB int limit (src/trie.cuh:1115). */
int fm_graph_save_bin(const fm_graph *g, const char *path);
int fm_graph_load_bin(const char *path, fm_graph **out);
/* Writes the graph in the reference's .g grammar (paired edge lines as
 * stored; the file ends with "t # -1" and no newline, src/trie.cuh:1080). */
int fm_graph_write_g(const fm_graph *g, const char *path);
void fm_graph_destroy(fm_graph *g);

/* ---- synthetic data graphs (seeded, deterministic) ----------------------
 * Undirected simple graphs written as paired directed edges with the same
 * label on both lines. label_dist: 0 = uniform over [0, nlabels),
 * 1 = Zipf P(k) ~ (k+1)^-zipf_s over [0, nlabels). Vertex labels uniform
 * over [0, vlabels). */
int fm_gen_erdos_renyi(uint32_t vertices, uint64_t undirected_edges,
                       uint32_t nlabels, int label_dist, double zipf_s,
                       uint32_t vlabels, uint64_t seed, fm_graph **out);
int fm_gen_rmat(uint32_t scale, uint64_t undirected_edges, double a, double b,
                double c, uint32_t nlabels, int label_dist, double zipf_s,
                uint32_t vlabels, uint64_t seed, fm_graph **out);

/* Query extraction by seeded random walk on a symmetric data graph (the
 * walk crosses edges of every label). shape 0: tree (the edge that first
 * reached each vertex); 1: a weighted-growth tree plus min_cycles (or one
 * more) induced chords; 2: the whole induced subgraph of the grown vertex
 * set; 3: the growth tree plus exactly min_cycles chords; 4: a cycle
 * through `vertices` vertices (a self-avoiding walk closed by an edge back
 * to its start; the query of the 1,000-vertex configuration is a 4-cycle).
 * Edge labels are copied from the data, so the query has at least one
 * embedding. */
int fm_gen_query_walk(const fm_graph *g, uint32_t vertices, int shape,
                      uint32_t min_cycles, uint64_t seed, fm_query **out);

/* ---- queries ------------------------------------------------------------
 * Replaces fileParse (src/hostparser.cuh:286-509) and preProcessQuery
 * (src/ccsr.cuh:485-728). Query edge lines are directed; an embedding m
 * must map every line (u, w, l) to a data edge (m(u), m(w)) of label l,
 * with m injective. Vertex labels are ignored. */
int fm_query_load_g(const char *path, fm_query **out);
int fm_query_from_edges(uint32_t vertices, uint32_t nedges,
                        const uint32_t *src, const uint32_t *dst,
                        const uint32_t *label, fm_query **out);
int fm_query_write_g(const fm_query *q, const char *path);
int fm_query_info(const fm_query *q, uint32_t *vertices, uint32_t *nedges);
/* Copies the unique edge lines (nedges entries each). */
int fm_query_edges(const fm_query *q, uint32_t *src, uint32_t *dst,
                   uint32_t *label);
/* Matching plan, following the reference's order (src/ccsr.cuh:229-249):
 * vertices ranked by (out-degree desc, id asc); depth 0 = rank 0; then the
 * lowest-ranked unmapped neighbour of the mapped set.
 * order[d] = query vertex matched at depth d. For each depth d,
 * req_begin[d] .. req_begin[d+1] index the requirements; requirement i is
 * (req_anchor[i] = depth of the earlier vertex, or d itself for a
 * self-loop; req_label[i] = label value; req_dir[i] = 0 if the candidate
 * must be in the anchor's out-row (query line anchor -> u), 1 if in its
 * in-row (query line u -> anchor)). Arrays must hold vertices entries
 * (order), vertices + 1 (req_begin) and nreqs entries. */
typedef struct fm_plan_info {
    uint32_t vertices;
    uint32_t nreqs;
    uint32_t max_reqs_per_depth;
} fm_plan_info;
int fm_query_plan(const fm_query *q, fm_plan_info *info, uint32_t *order,
                  uint32_t *req_begin, uint32_t *req_anchor,
                  uint32_t *req_label, uint32_t *req_dir);
/* Automorphisms of q and the symmetry-breaking constraints used by
 * opts.symmetry: auts (n_aut x vertices entries, auts[k*n + u] = p_k(u),
 * the identity first) and less (2 x n_less entries: m(less[2i]) <
 * m(less[2i+1])). Arrays may be NULL to query the sizes; when q has more
 * than max_auts automorphisms, *n_aut = 0 and none are listed. */
#define FM_MAX_AUTOMORPHISMS 48
int fm_query_symmetry(const fm_query *q, uint32_t max_auts, uint32_t *n_aut,
                      uint32_t *auts, uint32_t aut_cap, uint32_t *n_less,
                      uint32_t *less, uint32_t less_cap);
/* The plan fm_match uses for q on graph g under `plan` (FM_PLAN_*), in the
 * layout of fm_query_plan. */
int fm_query_plan_on(const fm_graph *g, const fm_query *q, uint32_t plan,
                     fm_plan_info *info, uint32_t *order, uint32_t *req_begin,
                     uint32_t *req_anchor, uint32_t *req_label,
                     uint32_t *req_dir);
/* In a table-mode launch (fm_match, fm_match_batch), count q's embeddings
 * without storing its rows (count_only = 1), as a count-mode launch does:
 * the reference's WriteSolnFlag = false (src/environment.cuh:65) for one
 * query of a batch, so that queries with too many embeddings for a table
 * run beside queries with tables. Its result has rows = 0 and no
 * FM_RESULT_TRUNCATED flag. No effect in a count-mode launch or a server. */
int fm_query_set_count_only(fm_query *q, uint32_t count_only);
void fm_query_destroy(fm_query *q);

/* ---- engine -------------------------------------------------------------
 * Replaces prealloc()/populatePool and the MemoryPool/forest of
 * src/tbs.cuh:10-136 and src/smsm.cu:1155-1180,1230-1262. Zero fields take
 * the defaults of include/environment.h. */
typedef struct fm_engine_config {
    uint64_t soln_prealloc;   /* bank arena bytes (SOLN_PREALLOC) */
    uint64_t output_prealloc; /* table arena bytes (OUTPUT_PREALLOC) */
    uint32_t bank_size;       /* bank size in 8-byte units (CurBankSize) */
    uint32_t output_page_bytes;
    uint32_t queue_capacity;  /* task slots per priority level (pow2) */
    uint32_t max_queries;     /* queries per launch */
    int32_t device;           /* HIP device id (default 0) */
    uint32_t reserved0;
    /* PREALLOC (src/environment.cuh:4): device bytes that the engine's
     * arenas and the graph it matches on may take together (0 = PREALLOC of
     * include/environment.h; FM_PREALLOC_UNLIMITED = no limit). Beyond it,
     * fm_engine_create, fm_match* and fm_server_create return
     * FM_ERR_GENERIC (1), the reference's exit status when its bump
     * allocator overflows the slab (src/trie.cuh:1034-1049). */
    uint64_t prealloc;
    uint32_t reserved[4];
} fm_engine_config;
#define FM_PREALLOC_UNLIMITED (~0ull)
int fm_engine_create(const fm_engine_config *cfg, fm_engine **out);
void fm_engine_destroy(fm_engine *e);
/* Device and launch-shape facts: compute units, resident waves of the
 * matcher kernel, arena sizes. */
typedef struct fm_engine_info {
    uint32_t compute_units;
    uint32_t waves_per_cu;
    uint32_t grid_waves;
    uint32_t banks;
    uint64_t bank_bytes;
    uint32_t output_pages;
    uint32_t output_page_bytes;
    char device_name[64];
    uint64_t arena_bytes;     /* device bytes of the arenas, queues and
                                 counters (what PREALLOC counts besides the
                                 graph) */
} fm_engine_info;
int fm_engine_info_get(fm_engine *e, fm_engine_info *out);

/* ---- matching -------------------------------------------------------------
 * Replaces match() + initPosting + the dynamic-parallelism fork chain
 * (src/smsm.cu:784-1139,1182-1349): one persistent kernel launch whose
 * waves pull partial-embedding tasks from device work queues. */
#define FM_MODE_COUNT 0u  /* count only; the last level is never stored */
#define FM_NO_DEADLINE 0xFFFFFFFFu  /* fm_match_opts.timeout_ms: no deadline */
#define FM_MODE_TABLE 1u  /* count and store every embedding */

/* Matching order. The embedding set does not depend on it; the work does.
 * FM_PLAN_COST orders by estimated candidates from the data graph's label
 * statistics (no reference counterpart); FM_PLAN_REFERENCE is the
 * reference's query-degree order (src/ccsr.cuh:229-249). */
enum { FM_PLAN_COST = 0, FM_PLAN_REFERENCE = 1 };

typedef struct fm_match_opts {
    uint32_t mode;            /* FM_MODE_COUNT or FM_MODE_TABLE */
    uint32_t instrument;      /* 1: gather algorithmic-byte counters */
    uint64_t max_rows;        /* table capacity per query (0 = arena) */
    uint64_t max_count;       /* abort a query past this count (0 = none) */
    uint32_t timeout_ms;      /* launch deadline (0 = 120000; FM_NO_DEADLINE:
                                 none, as the reference's CLI runs) */
    uint32_t grid_waves;      /* persistent waves (0: 8 per CU for one query,
                                 of which 4 start at once and 4 join after
                                 FM_JOIN_MS; every resident wave for several;
                                 an explicit value starts all at once and is
                                 capped at the resident waves: 8 per CU for
                                 one query, 12 for several on gfx950) */
    uint32_t task_rows;       /* parents per queued task (0: FM_QUERY_TASK_ROWS for
                                 one query, FM_BATCH_TASK_ROWS for several) */
    uint32_t root_chunk;      /* roots per claim (0: TooLargeThreshold; one
                                 query: ~8 chunks per starting wave, 64 to
                                 TooLargeThreshold roots) */
    uint32_t plan;            /* FM_PLAN_COST (0) or FM_PLAN_REFERENCE */
    /* abort a query once it has stored more than this many intermediate
     * partial embeddings (0 = none); its status becomes COUNT_LIMIT */
    uint64_t max_partials;
    /* 1: exact symmetry breaking. Embeddings are enumerated once per orbit
     * of the query's automorphism group (ordering constraints on the
     * matched vertices) and every one found is expanded by the group, so the
     * count and the table are those of full enumeration. Queries with more
     * than FM_MAX_AUTOMORPHISMS automorphisms are enumerated in full. No
     * reference counterpart: the reference's rule (ccsr.cuh:342-676,
     * smsm.cu:187-227) is inexact for C5, C6 and the bowtie. */
    uint32_t symmetry;
    /* Waves a query's roots fan out to when it starts: its root list is cut
     * into this many claims (root_chunk is then ignored; 0: root_chunk
     * roots per claim). Deeper work spreads further by splitting. */
    uint32_t waves_per_query;
    /* LDS tile: parent rows staged per extension step, with their
     * requirement rows (1..64; 0: 64; wide rows may stage fewer, as the
     * pool holds 2,048 words). */
    uint32_t tile_parents;
    uint32_t reserved[1];
} fm_match_opts;

/* One query in one launch (the reference's match()). */
int fm_match(fm_engine *e, const fm_graph *g, const fm_query *q,
             const fm_match_opts *opts, fm_result **out);
/* nq queries in one persistent launch (no reference counterpart: the
 * reference matches one query per process). */
int fm_match_batch(fm_engine *e, const fm_graph *g, const fm_query *const *qs,
                   uint32_t nq, const fm_match_opts *opts, fm_result **out);

typedef struct fm_query_result {
    uint64_t count;       /* number of embeddings (exact unless ABORTED) */
    uint64_t rows;        /* rows held in the table (<= count) */
    uint32_t status;      /* fm_status of this query */
    uint32_t flags;       /* FM_RESULT_* */
    uint32_t vertices;    /* row width */
    uint32_t reserved;
    /* Instrumented launches only: algorithmic bytes (neighbour words read
     * once per requirement, 8-byte partial records, the final table) and
     * the number of row lookups. */
    uint64_t alg_bytes;
    uint64_t row_lookups;
    uint64_t partials[FM_MAX_QUERY_VERTICES_ABI];
    /* candidates drawn from driver rows, and membership probes into the
     * other requirement rows (instrumented launches only) */
    uint64_t candidates;
    uint64_t probes;
} fm_query_result;

typedef struct fm_launch_stats {
    double kernel_ms;        /* persistent kernel, HIP events on its stream */
    double total_ms;         /* host wall time of the call */
    uint64_t tasks_pushed;   /* queued tasks (forks) */
    uint64_t root_claims;
    uint64_t banks_allocated;/* fresh banks from the bump allocator */
    uint64_t banks_reused;   /* banks taken from the free list */
    uint64_t alloc_failures; /* requeues after a failed bank allocation */
    uint64_t pages_allocated;
    uint64_t tasks_split;    /* tasks donated to idle waves (work splitting) */
    uint64_t requeues;       /* tasks re-queued by a yielding wave */
    uint32_t grid_waves;
    uint32_t status;
    double busy_frac;        /* share of the waves' resident time spent in tasks */
} fm_launch_stats;

uint32_t fm_result_queries(const fm_result *r);
int fm_result_query(const fm_result *r, uint32_t qi, fm_query_result *out);
int fm_result_stats(const fm_result *r, fm_launch_stats *out);
/* Copies up to max_rows rows (row-major, `vertices` u32 columns in query
 * vertex id order) of query qi into dst. Returns the rows copied through
 * *copied. The order of rows is unspecified; sorted=1 sorts them
 * lexicographically on the host. */
int fm_result_copy_table(const fm_result *r, uint32_t qi, uint32_t *dst,
                         uint64_t max_rows, int sorted, uint64_t *copied);
void fm_result_destroy(fm_result *r);

/* ---- serving ---------------------------------------------------------------
 * No reference counterpart (the reference matches one query per process).
 * A server keeps the graph and the bank pool resident and runs one
 * persistent matcher launch that admits queries while others are running
 * (continuous batching): the host writes a query into a slot and appends it
 * to a submission ring in pinned memory; an idle wave admits it by queueing
 * its root chunks, so its work interleaves with the queries already on the
 * device. Completion (count, rows, status) is published to pinned memory.
 * The launch ends after `idle_ms` without work and is relaunched by the next
 * submit or poll. Admission is refused (FM_ERR_BUSY) when all slots are in
 * use or the live banks exceed `bank_budget`. */
#define FM_ERR_BUSY 311
typedef struct fm_server fm_server;
typedef struct fm_server_config {
    uint32_t slots;          /* concurrent queries (default 256) */
    uint32_t slot_pages;     /* output pages per query (default 256) */
    uint32_t bank_budget;    /* admit while live banks < budget (0 = pool) */
    uint32_t idle_ms;        /* launch lifetime without work (default 20) */
    uint32_t root_chunk;     /* roots per root task (default TooLargeThreshold) */
    uint32_t task_rows;      /* parents per task (default SubMaxMulSpill) */
    uint32_t table;          /* 1: store tables (default), 0: count only */
    uint32_t max_launch_ms;  /* deadline of one launch (default 120000) */
    uint32_t plan;           /* FM_PLAN_COST (0) or FM_PLAN_REFERENCE */
    uint32_t slice_ms;       /* a launch retires after this long and the next
                                submit/poll relaunches it; queued work stays
                                on the device (default 2000) */
    uint32_t symmetry;       /* as fm_match_opts.symmetry */
    uint32_t reserved[5];
} fm_server_config;
typedef struct fm_server_completion {
    uint32_t ticket;         /* returned by fm_server_submit */
    uint32_t status;
    uint32_t flags;
    uint32_t vertices;
    uint64_t count;
    uint64_t rows;
    double submit_ms;        /* host clock, ms since the server started */
    double done_ms;          /* host clock when the completion was seen */
} fm_server_completion;
typedef struct fm_server_stats {
    uint64_t submitted, completed, refused, launches;
    uint32_t in_flight, live_banks;
    uint64_t dropped_tasks;  /* tasks of already completed queries (should be 0) */
    uint64_t refused_memory; /* of `refused`: a slot was free but the live banks
                                had reached bank_budget (admission by memory) */
} fm_server_stats;
int fm_server_create(fm_engine *e, const fm_graph *g,
                     const fm_server_config *cfg, fm_server **out);
/* Queues a query (max_rows = table capacity, 0 = slot pages). The ticket
 * names the request until fm_server_release. */
int fm_server_submit(fm_server *s, const fm_query *q, uint64_t max_rows,
                     uint32_t *ticket);
/* Returns up to `max` completions not returned before; non-blocking. */
int fm_server_poll(fm_server *s, fm_server_completion *out, uint32_t max,
                   uint32_t *n);
/* Copies a completed request's table (query-vertex column order). */
int fm_server_copy_table(fm_server *s, uint32_t ticket, uint32_t *dst,
                         uint64_t max_rows, int sorted, uint64_t *copied);
/* Frees the request's slot (its pages are recycled when the slot is reused). */
int fm_server_release(fm_server *s, uint32_t ticket);
int fm_server_stats_get(fm_server *s, fm_server_stats *out);
/* Waits for the in-flight requests and stops the launch. */
void fm_server_destroy(fm_server *s);

/* ---- intersection microkernel -------------------------------------------
 * Replaces the warp-cooperative intersect<CoopSize,Decomp> family
 * (src/intersect.cuh:31-191). pairs of strictly increasing u32 lists inside
 * one device array `data` (device pointers): list A_i = data[a_off[i] ..
 * +a_len[i]], B_i likewise. counts[i] = |A_i n B_i|. If out_off is not NULL,
 * the intersection of pair i is written ascending at out[out_off[i] ..]
 * (room for min(a_len[i], b_len[i]) words). Lists with repeated values give
 * unspecified counts, never accesses outside the lists and that room. */
int fm_intersect_pairs(fm_engine *e, const uint32_t *data,
                       const uint64_t *a_off, const uint32_t *a_len,
                       const uint64_t *b_off, const uint32_t *b_len,
                       uint32_t npairs, uint32_t *counts,
                       const uint64_t *out_off, uint32_t *out,
                       double *kernel_ms);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc. fm_intersect_pairs with its work array (claim counters,
 * the pairs' first segments, one status word per segment) in a fixed device
 * array of the library (FM_NOALLOC_SCRATCH_BYTES per GPU, held until the
 * process ends) instead of engine memory; it calls no hipMalloc, hipFree or
 * other allocating routine (the engine, created beforehand, provides the
 * device and stream). Limits: FM_NOALLOC_MAX_PAIRS pairs, and the pairs'
 * segments (16,384 merge-path elements each, or 512 elements of the shorter
 * list of a pair whose longer list is 32x longer) must fit the array: 491,359
 * segments.
 * Beyond either limit it returns FM_ERR_INVALID_ARG (counts are then zero).
 * Every call on one GPU uses that GPU's one array: calls that overlap on one
 * GPU (from several engines or threads) race. */
#define FM_NOALLOC_MAX_PAIRS (1u << 16)
#define FM_NOALLOC_SCRATCH_BYTES (4u << 20)
int fm_intersect_pairs_noalloc(fm_engine *e, const uint32_t *data,
                               const uint64_t *a_off, const uint32_t *a_len,
                               const uint64_t *b_off, const uint32_t *b_len,
                               uint32_t npairs, uint32_t *counts,
                               const uint64_t *out_off, uint32_t *out,
                               double *kernel_ms);

/* Device buffers for the microkernel and its tests (hipMalloc/hipMemcpy). */
int fm_device_alloc(uint64_t bytes, void **out);
int fm_device_free(void *p);
int fm_copy_to_device(void *dst, const void *src, uint64_t bytes);
int fm_copy_to_host(void *dst, const void *src, uint64_t bytes);

#ifdef __cplusplus
}
#endif

#endif /* FORKMATCH_H */
