// ForkMatch <query.g> <data.g>
//
// Drop-in for the reference CLI (src/smsm.cu:1351-1393): the same argv,
// the same exit codes (99 for a wrong argument count, the parse codes of
// the loaders) and the same stdout lines "Soln Size", "Kernel Time",
// "End to End Time" and "Pop Count".
//
// Knobs: PREALLOC-family defaults come from environment.h; each may be
// overridden by an environment variable of the same name (PREALLOC,
// SOLN_PREALLOC, OUTPUT_PREALLOC, CurBankSize). PREALLOC bounds the arenas
// plus the graph; beyond it the CLI exits with status 1, as the reference
// does when its slab overflows (src/trie.cuh:1034-1049). The match runs to
// completion like the reference's (no deadline); FM_TIMEOUT_MS=<ms> sets
// one. FM_TABLE_OUT=<path> writes the sorted
// embedding table as text, one row per line, columns in query vertex order.
// FM_COUNT_ONLY=1 skips table construction (the reference's WriteSolnFlag,
// src/environment.cuh:65).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "environment.h"
#include "forkmatch.h"

static uint64_t env_u64(const char *name, uint64_t dflt) {
    const char *v = getenv(name);
    if (!v || !*v) return dflt;
    return strtoull(v, nullptr, 0);
}

static int die(int code) {
    fprintf(stdout, "Error: %s\n", fm_last_error());
    return code;
}

int main(int argc, char **argv) {
    printf("Trie Graph Problem\n");
    if (argc != 3) return FM_ERR_ARGV;
    auto start = std::chrono::steady_clock::now();
    const char *query_path = argv[1], *data_path = argv[2];
    printf("Query Graph - %s Data Graph - %s\n", query_path, data_path);
    // same order as the reference's match(): data graph, then query
    fm_graph *g = nullptr;
    // The data graph is parsed on the device, as in the reference. Without a
    // visible device the host parser checks the file, so the file errors
    // keep their exit codes (matching then stops at engine creation).
    int rc = fm_device_count() > 0 ? fm_graph_load_g(data_path, &g) : fm_graph_parse_g(data_path, &g);
    if (rc) return die(rc);
    fm_query *q = nullptr;
    rc = fm_query_load_g(query_path, &q);
    if (rc) return die(rc);
    fm_engine_config cfg{};
    cfg.soln_prealloc = env_u64("SOLN_PREALLOC", SOLN_PREALLOC);
    cfg.output_prealloc = env_u64("OUTPUT_PREALLOC", OUTPUT_PREALLOC);
    cfg.bank_size = (uint32_t)env_u64("CurBankSize", CurBankSize);
    cfg.prealloc = env_u64("PREALLOC", PREALLOC);
    fm_engine *e = nullptr;
    rc = fm_engine_create(&cfg, &e);
    if (rc) return die(rc > 255 ? 1 : rc);
    rc = fm_graph_upload(g);
    if (rc) return die(rc > 255 ? 1 : rc);
    auto parsed = std::chrono::steady_clock::now();
    printf("\nParse Time: %fs\n", std::chrono::duration<double>(parsed - start).count());
    fm_match_opts opts{};
    opts.mode = env_u64("FM_COUNT_ONLY", 0) ? FM_MODE_COUNT : FM_MODE_TABLE;
    opts.timeout_ms = (uint32_t)env_u64("FM_TIMEOUT_MS", FM_NO_DEADLINE);
    fm_result *r = nullptr;
    rc = fm_match(e, g, q, &opts, &r);
    if (rc && !r) return die(rc > 255 ? 1 : rc);
    fm_query_result qr{};
    fm_result_query(r, 0, &qr);
    fm_launch_stats st{};
    fm_result_stats(r, &st);
    printf("Kernel Time: %fs\n", st.kernel_ms / 1000.0);
    printf("Soln Size: %llu\n", (unsigned long long)qr.count);
    if (qr.flags & FM_RESULT_TRUNCATED) printf("Table truncated at %llu rows\n", (unsigned long long)qr.rows);
    const char *table_out = getenv("FM_TABLE_OUT");
    if (table_out && *table_out && qr.rows) {
        std::vector<uint32_t> rows((size_t)qr.rows * qr.vertices);
        uint64_t copied = 0;
        if (fm_result_copy_table(r, 0, rows.data(), qr.rows, 1, &copied) == FM_OK) {
            FILE *f = fopen(table_out, "w");
            if (f) {
                for (uint64_t i = 0; i < copied; i++) {
                    for (uint32_t c = 0; c < qr.vertices; c++)
                        fprintf(f, c ? " %u" : "%u", rows[i * qr.vertices + c]);
                    fputc('\n', f);
                }
                fclose(f);
            }
        }
    }
    printf("Pop Count: %llu\n", (unsigned long long)(st.banks_allocated + st.banks_reused));
    int status = rc ? rc : (int)qr.status;
    fm_result_destroy(r);
    fm_query_destroy(q);
    fm_graph_destroy(g);
    fm_engine_destroy(e);
    auto end = std::chrono::steady_clock::now();
    printf("\nEnd to End Time: %fs\n", std::chrono::duration<double>(end - start).count());
    if (status) {
        fprintf(stdout, "Error: %s\n", fm_last_error());
        return status > 255 ? 1 : status;
    }
    return 0;
}
