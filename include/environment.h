/*
 * Build-time defaults for the ForkMatch engine knobs.
 *
 * The names are the reference's compile-time knobs
 * (src/environment.cuh:4-20, README.md:37-49) so that a user who tuned them
 * there finds them here. Unlike the reference they are only defaults: every
 * one of them can be overridden at run time through fm_engine_config
 * (include/forkmatch.h) or, for the CLI, through an environment variable of
 * the same name (e.g. `CurBankSize=131072 ForkMatch q.g d.g`).
 */
#pragma once

/* Total device memory the CLI reserves for one match (src/environment.cuh:4).
 * Here it is the sum of the two arenas below plus the graph; it is not one
 * slab, each arena is allocated on its own. */
#define PREALLOC 45600000000ull

/* Bytes of the partial-embedding bank arena (src/environment.cuh:5). Banks of
 * CurBankSize * 8 bytes are carved from it by a device bump allocator and
 * recycled through a device free list. */
#define SOLN_PREALLOC 24000000000ull

/* Bytes of the output-table arena (src/environment.cuh:6). Tables are stored
 * in lazily allocated pages of OUTPUT_PAGE_BYTES. */
#define OUTPUT_PREALLOC 17300000000ull

/* Size of one memory bank in 8-byte units, the width of the reference's
 * FuncPair node (src/environment.cuh:19). The reference default is 1 << 20
 * (8 MiB banks). Here a wave yields its task whenever its output bank
 * fills, so the bank is also the unit of depth-first progress; the default
 * is 1 << 15 (256 KiB banks, 96k banks in SOLN_PREALLOC). The minimum is
 * 2048 (one step of 2 x 64 rows of 32 vertices). */
#define CurBankSize (1u << 15)

/* Threads per workgroup of the reference's main kernel
 * (src/environment.cuh:11). The MI355X engine runs one wave64 per
 * workgroup; the name is kept for the launch-shape report only. */
#define CurBlockSize 64

/* Maximum parents carried by one queued task. This replaces the reference's
 * soft fork threshold (outputMax - BlockSize * SubMaxMulSpill,
 * src/smsm.cu:451-454): a bank segment larger than this is split into
 * several tasks so idle waves can take part of it. While waves are idle a
 * wave cuts its children into 64-parent tasks instead, and donates half of
 * its remaining parents or candidates. Measured on R-MAT 2^20 with 16
 * queries of 6 vertices, one launch each: 1024-parent tasks take 100 ms per
 * step, 64-parent tasks 215 ms (per-task queue and staging overhead, and
 * siblings of one parent kept together share their neighbour rows in L2).
 * The serving loop uses it; one-query launches use FM_QUERY_TASK_ROWS. */
#define SubMaxMulSpill 1024

/* The same for a launch that carries one query. On R-MAT 2^20 with the 32
 * plain queries of 6-8 vertices (16 counted without tables), one launch
 * each: 1024-parent tasks 259 ms of kernel time per step, 2048-parent tasks
 * 243 ms, 4096-parent tasks 245 ms; the count-only trees gain 18-27% (every
 * task costs several atomics on the queue, pending and bank words that all
 * waves share). */
#ifndef FM_QUERY_TASK_ROWS
#define FM_QUERY_TASK_ROWS 2048
#endif

/* The same for launches that carry several queries: there the queries
 * themselves keep the waves busy. Measured on 1,024 queries of 4-12 vertices
 * (R-MAT 2^20) at 12 waves per CU: 4096-parent tasks 345 ms, 1024-parent
 * tasks 380 ms; 8192 and 16384 within 2% of 4096. */
#define FM_BATCH_TASK_ROWS 4096

/* A launch of one query with the default grid starts half of its waves
 * (4 per CU) and lets the other half join once it has run this many
 * milliseconds: short launches lose more to per-wave start costs and shared
 * counters than extra waves give back, long ones gain from them. Measured
 * on R-MAT 2^20 with 16 queries of 6 vertices: the one 30 ms query runs in
 * 22-24 ms with a join after 5-9 ms, while queries of 6-7 ms run 10-15%
 * slower when the waves join after 5 ms. */
#ifndef FM_JOIN_MS
#define FM_JOIN_MS 8
#endif

/* Number of root vertices claimed by one wave at a time (the reference gives
 * each block a fixed root range, src/smsm.cu:803-811). A one-query launch
 * with the default takes about 8 chunks per starting wave instead, between
 * 64 roots and this many. */
#define TooLargeThreshold 1024

/* Bytes of one output-table page. A wave appends to its own page; a
 * partly filled page is handed to the next wave of the same query. */
#define OUTPUT_PAGE_BYTES (1u << 18)

/* Limits of the device plan. The reference supports queries of at most 19
 * vertices and 3 backward edges per vertex (src/ccsr.cuh:381-385,
 * src/smsm.cu:887-899); here the limits are 32 vertices and 16 backward
 * edges per vertex. */
#define FM_MAX_QUERY_VERTICES 32
#define FM_MAX_REQS_PER_DEPTH 16
#define FM_MAX_REQS_TOTAL 256
