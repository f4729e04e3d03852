"""ForkMatch MI355X benchmark.

Default workload (N=1): plain matching at scale, with the batched form
beside it. An R-MAT data graph (a, b, c, d = .57, .19, .19, .05), 2^20
vertices, 2^24 undirected edges written as 2^25 directed edges, edge labels
Zipf(1.2) over 0..86, resident in HBM.
 - Plain (the headline): 32 seeded random-walk queries of 6-8 vertices, each
   matched in its own launch. 16 are tree-like (a walk tree or a tree plus
   one chord), enumerated in count mode (the reference's WriteSolnFlag =
   false, src/environment.cuh:65) and screened to at most 2^31 embeddings;
   16 have >= 2 independent cycles and are matched with their full embedding
   table written to HBM (at most 2^26 embeddings each). One step = all 32 launches. value = embeddings/s over the timed
   steps (whole job, inputs resident before the timed region).
 - Batched (the `batched` object): 1,024 queries of 4..12 vertices (shape
   caps 40% trees, 40% cycle+chord, 20% near-cliques, each with a table of at
   most 2^22 rows) in ONE persistent launch, against the same queries one
   launch each: queries/s of both and their ratio.

--config 3: the batched leg alone (value = batched queries/s).

--config 5: stress. R-MAT 2^24 vertices, 2^28 undirected edges, edge labels
uniform over 0..3; 64 dense 10-vertex queries (15 edges each), each capped at
2.5e9 / 64 rows (~100 GB of tables in all), matched in one batched launch
with every table resident; value = rows stored per second.

--config 4: serving. R-MAT 2^22 vertices, 2^26 undirected edges, edge
labels Zipf(1.2) over 0..127, 4 vertex labels; 8 MiB banks (CurBankSize =
1 << 20). 512 queries of 4..16 vertices (at most 2^18 embeddings each)
form the request pool; Poisson arrival streams at 200, 1,000 and 5,000
requests/s go through one continuous-batching server, admission by bank
budget. value = completed queries/s at the highest offered rate; p50/p99
latency (arrival to completion seen by the client) per rate.

Prints one JSON line. Only rank 0 prints; --gpus above 1 is refused (the
build targets one MI355X).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import forkmatch_amd as fm  # noqa: E402

HBM_PEAK_GBS = 8000.0
FM_ERR_COUNT_LIMIT = 306  # a query passed max_count (include/forkmatch.h)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_graph(cfg: int, seed: int):
    if cfg in (2, 3):
        return fm.Graph.rmat(20, 1 << 24, 87, zipf_s=1.2, seed=seed, vlabels=2)
    if cfg == 4:
        return fm.Graph.rmat(22, 1 << 26, 128, zipf_s=1.2, seed=seed, vlabels=4)
    if cfg == 5:
        return fm.Graph.rmat(24, 1 << 28, 4, zipf_s=0.0, seed=seed, vlabels=8)
    raise SystemExit(f"unsupported config {cfg}")


def make_candidates(g, cfg: int, seed: int, count: int):
    """Seeded candidate queries: (shape, n, Query)."""
    rng = np.random.default_rng(seed)
    out = []
    i = 0
    while len(out) < count:
        i += 1
        if cfg == 5:
            # dense: 10 vertices, 15 edges (a spanning tree plus 6 chords, the
            # cyclomatic number of a 10-cycle with 5 chords)
            try:
                q = fm.Query.random_walk(g, 10, shape=3, min_cycles=6, seed=seed * 1_000_003 + i)
            except fm.ForkMatchError:
                continue
            out.append((3, 10, q))
            continue
        if cfg == 2:
            # half tree-like (at most one cycle: a walk tree, or the growth
            # tree plus one chord; matched count-only), half with >= 2
            # independent cycles (matched with tables)
            kind = len(out) % 4
            shape = (0, 1, 3, 1)[kind]
            n = int(rng.integers(6, 9))
            cyc = 2 if shape == 1 else (1 if shape == 3 else 0)
        else:
            u = rng.random()
            shape = 0 if u < 0.4 else (1 if u < 0.8 else 2)
            n = int(rng.integers(4, 13 if cfg == 3 else 17))
            cyc = 1 if shape == 1 else 0
        try:
            q = fm.Query.random_walk(g, n, shape=shape, min_cycles=cyc, seed=seed * 1_000_003 + i)
        except fm.ForkMatchError:
            continue
        out.append((shape, n, q))
        if len(out) % 256 == 0:
            log(f"candidates: {len(out)}/{count}")
    return out


def screen(engine, g, cands, budget: int, want, per_shape=None, max_partials=None, total_cap=None):
    """Keeps candidates whose count is in [1, budget], screened in count mode
    in batched launches (untimed). want = total wanted; per_shape caps;
    total_cap bounds the sum of the kept counts."""
    keep, counts = [], []
    taken = {}
    hist = {}
    for b0 in range(0, len(cands), 64):
        chunk = cands[b0:b0 + 64]
        r = engine.match_batch(g, [c[2] for c in chunk], table=False, max_count=budget,
                               max_partials=budget if max_partials is None else max_partials, timeout_ms=120000,
                               allow_errors=True)
        for j, c in enumerate(chunk):
            o = r.query(j)
            hist[o.status] = hist.get(o.status, 0) + 1
            if o.status != 0 or o.count == 0 or o.count > budget:
                continue
            if per_shape is not None and taken.get(c[0], 0) >= per_shape.get(c[0], 0):
                continue
            if total_cap is not None and sum(counts) + o.count > total_cap:
                continue
            taken[c[0]] = taken.get(c[0], 0) + 1
            keep.append(c)
            counts.append(o.count)
            if len(keep) >= want:
                log(f"screen: statuses {hist}")
                return keep, counts
        r.close()
        log(f"screen: {b0 + len(chunk)}/{len(cands)} screened, {len(keep)} kept")
    log(f"screen: statuses {hist}")
    return keep, counts


def select_mixed(engine, g, cfg: int, seed: int, nq: int, budget: int):
    """Progressive screening from a seeded pool (at most 48 x nq candidates)
    with the 40/40/20 tree / cycle+chord / near-clique caps; shapes that
    cannot fill their share within the pool are topped up by the others."""
    caps = {0: int(0.4 * nq), 1: int(0.4 * nq), 2: nq - 2 * int(0.4 * nq)}
    chosen, counts, pool = [], [], []
    rnd = 0
    while len(chosen) < nq and len(pool) < 48 * nq:
        cands = make_candidates(g, cfg, seed + 7919 * rnd, max(64, nq // 2))
        rnd += 1
        pool += cands
        left = {k: caps[k] - sum(1 for c in chosen if c[0] == k) for k in caps}
        c2, n2 = screen(engine, g, cands, budget, nq - len(chosen), per_shape=left)
        chosen += c2
        counts += n2
    if len(chosen) < nq:
        taken = {id(c[2]) for c in chosen}
        c2, n2 = screen(engine, g, [c for c in pool if id(c[2]) not in taken], budget, nq - len(chosen))
        chosen += c2
        counts += n2
    return chosen, counts


_CSR = {}


def validate_sample(g, q, table: np.ndarray, rng, k=256) -> bool:
    """Every sampled row is an injective map satisfying every query edge."""
    if len(table) == 0:
        return True
    if id(g) not in _CSR:
        _CSR[id(g)] = g.csr(0)
    ptr, col, labels = _CSR[id(g)]
    V = g.vertices
    lab_index = {int(v): i for i, v in enumerate(labels)}
    edges = q.edges()
    rows = table[rng.integers(0, len(table), min(k, len(table)))]
    for row in rows:
        if len(set(row.tolist())) != len(row):
            return False
        for s, d, l in edges:
            li = lab_index.get(int(l))
            if li is None:
                return False
            r = li * V + int(row[s])
            seg = col[ptr[r]:ptr[r + 1]]
            i = np.searchsorted(seg, row[d])
            if i >= len(seg) or seg[i] != row[d]:
                return False
    return True


def host_threads() -> tuple[int, int, int]:
    """(threads for the all-cores leg, CPUs this process may run on,
    os.cpu_count()). The pool is the process's CPU affinity, capped at
    OMP_NUM_THREADS when that is set: a GPU box shares its host among its
    GPUs and sets it to this GPU's share of the cores."""
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cap = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return (min(avail, cap) if cap > 0 else avail), avail, (os.cpu_count() or avail)


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baselines(g, chosen, target_s: float):
    """The oracle on all host threads of this GPU's share and on one thread,
    each on its own bounded root sample of the same queries (count only,
    the reference's span: enumeration without graph load or planning)."""
    import oracle as orc
    t0 = time.perf_counter()
    s, d, l = g.edge_list()
    og = orc.OracleGraph(g.vertices, s, d, l)
    log(f"cpu baseline: oracle graph built in {time.perf_counter() - t0:.1f} s")
    threads, avail, ncpu = host_threads()
    out = {}
    for leg, nt in (("all_cores", threads), ("one_thread", 1)):
        cnt, secs, stride = cpu_baseline(og, g.vertices, chosen, target_s, nt)
        out[leg] = {"value": cnt / secs if secs > 0 else None, "cores": nt, "embeddings": cnt,
                    "seconds": secs, "root_stride": stride}
        log(f"cpu baseline ({leg}, {nt} threads): {out[leg]['value']:.3g} embeddings/s")
    main = out["all_cores"]
    return {"value": main["value"], "unit": "embeddings/s", "cores": threads, "kind": "port",
            "sample": f"oracle (C backtracking, count only) on the {len(chosen)} queries, roots "
                      f"{main['root_stride'] // 2}, {main['root_stride'] + main['root_stride'] // 2}, ... "
                      f"({main['embeddings']} embeddings in {main['seconds']:.1f} s)",
            "one_thread": out["one_thread"], "all_cores": out["all_cores"],
            "host": {"cpu_model": cpu_model(), "os_cpu_count": ncpu, "affinity_cpus": avail,
                     "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}}


def cpu_baseline(og, V, queries, target_s: float, threads: int):
    """The oracle (C backtracking) on a strided root sample of the same queries:
    roots stride/2, 3 stride/2, ... (the midpoint of each stride window, so
    the sample is not dominated by vertex 0, R-MAT's largest hub). Each
    query's enumeration also stops after its share of twice the target, so
    one root with a huge subtree cannot stretch the run; the count of a
    stopped enumeration is what it enumerated in its time."""
    import oracle as orc
    cap = max(0.25, 2.0 * target_s / max(1, len(queries)))

    def sample(stride):
        cnt, secs = 0, 0.0
        for _, _, q in queries:
            c, t, _ = orc.timed_count(og, q.vertices, q.edges(), threads, cap, roots=(stride // 2, V), stride=stride)
            cnt += c
            secs += t
        return cnt, secs

    # calibrate the stride on ever denser samples (1, 16, 256, ... roots per
    # query) until one takes a twentieth of the target, so the probe of a
    # heavy query set stays short; then scale the stride to the target
    stride = max(1, V)
    while True:
        probe = sample(stride)[1]
        log(f"cpu baseline: probe at stride {stride} took {probe:.2f} s")
        if probe >= target_s / 20 or stride == 1:
            break
        stride = max(1, stride // 16)
    if probe > 0:
        stride = int(max(1, min(V, stride * probe / target_s)))
    for _ in range(4):
        total_cnt, total_s = sample(stride)
        log(f"cpu baseline: stride {stride}: {total_cnt} embeddings in {total_s:.2f} s")
        # thread start-up dominates a tiny probe: refine until the sample is
        # at least a third of the target
        if total_s >= target_s / 3 or stride == 1:
            break
        stride = max(1, int(stride * max(total_s, 1e-3) / target_s))
    return total_cnt, total_s, stride


def intersect_bench(engine, pairs: int = 2048, mean_len: int = 32768, reps: int = 5, seed: int = 7):
    """The wave64 intersection microkernel on synthetic sorted lists.
    Bytes per pair = 4 (|A| + |B|) + 4 |A n B| (lists read once, the
    intersection written once)."""
    rng = np.random.default_rng(seed)
    la = rng.integers(mean_len // 2, 3 * mean_len // 2, pairs)
    lb = rng.integers(mean_len // 2, 3 * mean_len // 2, pairs)
    lens = np.empty(2 * pairs, np.int64)
    lens[0::2], lens[1::2] = la, lb
    gaps = rng.integers(1, 5, int(lens.sum())).astype(np.uint32)
    offs = np.concatenate([[0], np.cumsum(lens)])
    data = np.cumsum(gaps, dtype=np.uint64)
    # restart every list at a random base so each list is sorted on its own
    starts = np.repeat(data[offs[:-1]] - rng.integers(0, 1000, 2 * pairs).astype(np.uint64), lens)
    data = (data - starts).astype(np.uint32)
    a_off, b_off = offs[0:-1:2].astype(np.uint64), offs[1::2].astype(np.uint64)
    out_off = np.concatenate([[0], np.cumsum(np.minimum(la, lb))]).astype(np.uint64)
    res = fm.intersect_pairs(engine, data, a_off, la.astype(np.uint32), b_off, lb.astype(np.uint32),
                             out_off[:-1], int(out_off[-1]))
    bufs = res["buffers"]
    inter = int(res["counts"].astype(np.int64).sum())
    # spot-check 8 pairs against numpy
    ok = True
    for i in rng.integers(0, pairs, 8):
        A = data[int(a_off[i]):int(a_off[i]) + int(la[i])]
        B = data[int(b_off[i]):int(b_off[i]) + int(lb[i])]
        ok = ok and int(res["counts"][i]) == len(np.intersect1d(A, B))
    ms = []
    for _ in range(reps):
        r = fm.intersect_pairs(engine, None, None, la.astype(np.uint32), None, None, out_off[:-1],
                               int(out_off[-1]), buffers=bufs)
        ms.append(r["kernel_ms"])
    nbytes = 4 * (int(la.sum()) + int(lb.sum())) + 4 * inter
    gbs = nbytes / (np.mean(ms) / 1e3) / 1e9
    out = {"kernel": "fm_isx_plan_kernel + fm_isx_kernel", "pairs": pairs, "elements": int(lens.sum()),
           "bytes": nbytes, "kernel_ms": float(np.mean(ms)), "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": gbs / HBM_PEAK_GBS, "checked": bool(ok), "traffic": None}
    # HBM bytes per launch of both kernels, measured with rocprofv3 FETCH_SIZE /
    # WRITE_SIZE on this workload and calibrated against a copy of known size
    tpath = ROOT / "profiles" / "fm_isx_kernel_intersect_2048pairs_32k_traffic.json"
    if tpath.exists() and pairs == 2048 and mean_len == 32768 and seed == 7:
        tr = json.loads(tpath.read_text())
        if tr.get("algorithmic_bytes_per_launch") == nbytes and tr.get("kernel") == "fm_isx_kernel":
            out["traffic"] = tr["hbm_bytes_per_launch"]
            out["traffic_source"] = str(tpath.relative_to(ROOT))
    return out


def serve_rate(srv, pool, counts, rate: float, seconds: float, rng, balg=None, deadline_s: float = 240.0):
    """One Poisson arrival stream through the server. Requests are submitted
    at their arrival time (or, when admission refuses them, as soon as a
    slot and bank budget free up: the wait counts toward their latency).
    Every completion's count is checked against the screening count."""
    from collections import deque
    n_req = max(1, int(rate * seconds))
    sched = np.cumsum(rng.exponential(1.0 / rate, n_req))
    qidx = rng.integers(0, len(pool), n_req)
    done = np.full(n_req, np.nan)
    t2req, backlog = {}, deque()
    nxt = ndone = bad = retries = 0
    inflight_max = 0
    emb = 0
    t0 = time.perf_counter()
    last_log = t0
    while ndone < n_req:
        now = time.perf_counter() - t0
        while nxt < n_req and sched[nxt] <= now:
            backlog.append(nxt)
            nxt += 1
        while backlog:
            k = backlog[0]
            t = srv.submit(pool[qidx[k]])
            if t is None:
                retries += 1
                break
            t2req[t] = k
            backlog.popleft()
        inflight_max = max(inflight_max, len(t2req))
        comps = srv.poll()
        if comps:
            tn = time.perf_counter() - t0
            for c in comps:
                k = t2req.pop(c.ticket)
                done[k] = tn
                want = counts[qidx[k]]
                if c.status != 0 or c.count != want:
                    bad += 1
                emb += c.count
                srv.release(c.ticket)
                ndone += 1
        tn = time.perf_counter()
        if tn - last_log > 20:
            last_log = tn
            st = srv.stats()
            log(f"serve {rate:.0f} req/s: {ndone}/{n_req} done, {len(t2req)} in flight, {len(backlog)} waiting; "
                f"live banks {st.live_banks}, launches {st.launches}, refused {st.refused}")
        if tn - t0 > deadline_s:
            raise RuntimeError(f"serving run at {rate} req/s did not drain in {deadline_s} s "
                               f"({ndone}/{n_req} done)")
    lat = (done - sched) * 1e3
    span = float(np.nanmax(done) - sched[0])
    slow = np.argsort(lat)[::-1][:5]
    out = {"offered_req_per_s": rate, "requests": n_req, "achieved_queries_per_s": n_req / span,
           "latency_ms_p50": float(np.percentile(lat, 50)), "latency_ms_p99": float(np.percentile(lat, 99)),
           "latency_ms_max": float(lat.max()), "embeddings_per_s": emb / span, "max_in_flight": inflight_max,
           "refused_submits": retries, "mismatches": bad, "seconds": span,
           "slowest": [{"arrival_s": round(float(sched[k]), 4), "latency_ms": round(float(lat[k]), 2),
                        "query": int(qidx[k])} for k in slow]}
    if balg is not None:
        out["alg_bytes"] = int(sum(balg[i] for i in qidx))
        out["alg_gbs_wall"] = out["alg_bytes"] / span / 1e9
    return out


def serve_mix_check(srv, pool, alone, rng, extra: int = 56) -> bool:
    """The first len(alone) pool queries submitted together with `extra`
    others: each table must equal the table of the same query matched alone."""
    import faulthandler
    reqs = list(range(len(alone))) + [int(i) for i in rng.integers(0, len(pool), extra)]
    rng.shuffle(reqs)
    t2q, got = {}, {}
    pending = list(reqs)
    t0 = last_log = time.perf_counter()
    log(f"mix check: {len(reqs)} requests, {len(alone)} tables compared")
    # a call that blocks inside the library shows up as a stack dump
    faulthandler.dump_traceback_later(100, exit=True)
    try:
        while pending or t2q:
            while pending:
                t = srv.submit(pool[pending[0]])
                if t is None:
                    break
                t2q[t] = pending.pop(0)
            for c in srv.poll():
                qi = t2q.pop(c.ticket)
                if qi < len(alone) and qi not in got:
                    got[qi] = (c.status, c.count, srv.table(c.ticket, c.rows, c.vertices, sorted=True), c.flags)
                srv.release(c.ticket)
            tn = time.perf_counter()
            if tn - last_log > 10:
                last_log = tn
                st = srv.stats()
                log(f"mix check: {len(reqs) - len(pending) - len(t2q)}/{len(reqs)} done, {len(t2q)} in flight "
                    f"(tickets {sorted(t2q)[:8]}), live banks {st.live_banks}, launches {st.launches}, "
                    f"dropped {st.dropped_tasks}")
            if tn - t0 > 60:
                raise RuntimeError(f"mix check did not drain in 60 s ({len(t2q)} in flight)")
    finally:
        faulthandler.cancel_dump_traceback_later()
    log(f"mix check: drained in {time.perf_counter() - t0:.1f} s")
    ok = True
    for qi, (count, table) in enumerate(alone):
        st, cnt, tab, flags = got[qi]
        same = st == 0 and cnt == count and tab.shape == table.shape and bool(np.array_equal(tab, table))
        if not same:
            log(f"mix check: query {qi}: status {st} flags {flags:#x} count {cnt} (alone {count}), "
                f"rows {tab.shape[0]} (alone {table.shape[0]})")
        ok = ok and same
    return ok


def b_alg(o, n: int, table: bool) -> int:
    """Algorithmic bytes: neighbour words of every requirement row of every
    partial embedding, an 8-byte record per stored partial, the final table.
    A count-only query stores no final level: no record or row for it."""
    records = sum(o.partials[:n]) if table else sum(o.partials[:max(n - 1, 0)])
    return o.alg_row_bytes + 8 * records + (4 * n * o.count if table else 0)


# A device-memory budget the benchmarks' explicitly sized arenas fit in
# (MI355X: 288 GB of HBM3E).
BENCH_PREALLOC = 270 * 10 ** 9


def select_plain(engine, g, seed: int, nq: int, budget: int, tree_budget: int):
    """Half tree-like queries (at most one cycle; count mode, at most
    tree_budget embeddings), half with >= 2 independent cycles (tables, at
    most budget), from seeded rounds of candidates. Returns (chosen, counts,
    tables)."""
    half = nq - nq // 2
    trees, tcounts, cyc, ccounts = [], [], [], []
    # on the 2^20-vertex graph about 1 walk tree of 6-8 vertices in 300 has
    # at most 2^31 embeddings (most unicyclic walks do), and 1 query with
    # >= 2 cycles in 30 at most 2^26
    for rnd in range(24):
        if len(trees) >= half and len(cyc) >= nq - half:
            break
        cands = make_candidates(g, 2, seed + 7919 * rnd, 4 * nq)
        if len(trees) < half:
            k, c = screen(engine, g, [x for x in cands if x[0] in (0, 3)], tree_budget, half - len(trees),
                          max_partials=1 << 28)
            trees += k
            tcounts += c
        if len(cyc) < nq - half:
            k, c = screen(engine, g, [x for x in cands if x[0] == 1], budget, nq - half - len(cyc))
            cyc += k
            ccounts += c
    return trees + cyc, tcounts + ccounts, [False] * len(trees) + [True] * len(cyc)


def run_plain(engine, g, queries, tables, counts, steps: int, warmup: int, validate: bool):
    """One launch per query; a step is every query once."""
    for _ in range(warmup):
        for q, t in zip(queries, tables):
            engine.match(g, q, table=t, timeout_ms=300000).close()
    step_times, kernel_ms, lat, busy = [], np.zeros(len(queries)), [], []
    splits = requeues = tasks = 0
    last = []
    for _ in range(steps):
        ts = time.perf_counter()
        last = []
        for i, (q, t) in enumerate(zip(queries, tables)):
            tq = time.perf_counter()
            r = engine.match(g, q, table=t, timeout_ms=300000)
            lat.append((time.perf_counter() - tq) * 1e3)
            st = r.stats()
            kernel_ms[i] += st.kernel_ms
            busy.append(st.busy_frac * st.kernel_ms)
            splits += st.tasks_split
            requeues += st.requeues
            tasks += st.tasks_pushed
            last.append(r.query(0))
            r.close()
        step_times.append(time.perf_counter() - ts)
    kernel_ms /= steps
    ok = all(o.status == 0 and o.count == counts[i] and o.rows == (o.count if tables[i] else 0)
             for i, o in enumerate(last))
    # table validity on a sample (untimed)
    rng = np.random.default_rng(1)
    for i in range(0, len(queries) if validate else 0, 4):
        if tables[i]:
            r = engine.match(g, queries[i], timeout_ms=300000)
            ok = ok and validate_sample(g, queries[i], r.table(0, sorted=False, max_rows=1 << 16), rng)
            r.close()
    ms_step = 1e3 * float(np.mean(step_times))

    def half(sel):
        idx = [i for i in range(len(queries)) if sel(i)]
        emb = sum(last[i].count for i in idx)
        kms = float(sum(kernel_ms[i] for i in idx))
        return {"queries": len(idx), "embeddings_per_step": emb, "kernel_ms_per_step": kms,
                "kernel_embeddings_per_s": emb / (kms / 1e3) if kms > 0 else None}

    return {"ms_per_step": ms_step, "kernel_ms": kernel_ms, "outcomes": last,
            "embeddings_per_step": sum(o.count for o in last),
            "latency_ms_p50": float(np.percentile(lat, 50)), "latency_ms_p99": float(np.percentile(lat, 99)),
            "wave_busy_frac": float(sum(busy) / (float(np.sum(kernel_ms)) * steps)),
            "per_step": {"tasks": tasks / steps, "tasks_split": splits / steps, "requeues": requeues / steps},
            "count_only_trees": half(lambda i: not tables[i]), "cyclic_tables": half(lambda i: tables[i]),
            "parity_checks_passed": bool(ok)}


def run_batched(g, seed: int, nq: int, budget: int, steps: int, warmup: int, save: str = "", load: str = ""):
    """The batched form: nq queries of 4..12 vertices in one persistent
    launch (knobs from the autotuner) against the same queries one launch
    each. Returns the leg's results and its queries. save/load: the query
    set, its counts, algorithmic bytes and tuned knobs (a loaded set runs
    no screening, instrumented or tuning launches)."""
    from forkmatch_amd.autotune import TuneResult, autotune
    # 1,024 tables resident at once: a larger table arena
    engine = fm.Engine(output_prealloc=96 << 30, prealloc=BENCH_PREALLOC)
    t1 = time.time()
    if load:
        man = json.loads((Path(load) / "manifest.json").read_text())
        chosen = [(m["shape"], m["n"], fm.Query.load_g(str(Path(load) / m["file"]))) for m in man["queries"]]
        counts = [m["count"] for m in man["queries"]]
        balg = [m["balg"] for m in man["queries"]]
        queries = [c[2] for c in chosen]
        tune = TuneResult(man["best"], man["best_ms"], [])
    else:
        chosen, counts = select_mixed(engine, g, 3, seed, nq, budget)
        queries = [c[2] for c in chosen]
        log(f"batched leg: {len(queries)} queries selected in {time.time() - t1:.1f} s")
        ri = engine.match_batch(g, queries, table=False, instrument=True, timeout_ms=300000)
        balg = [b_alg(ri.query(i), chosen[i][1], table=True) for i in range(len(queries))]
        ri.close()
        # 1,024 queries: 12 waves per CU (grid_frac 1.5) was faster than 8 in
        # every earlier sweep, and root chunks of 1,024 roots tie with 16-64
        # waves per query, so the sweep covers the other knobs
        space = {"waves_per_query": [16, 64], "tile_parents": [32, 64], "task_rows": [1024, 4096],
                 "grid_frac": [1.5]}
        tune = autotune(engine, g, queries, space=space, table=False, log=log)
        if save:
            d = Path(save)
            d.mkdir(parents=True, exist_ok=True)
            items = []
            for i, (shape, n, q) in enumerate(chosen):
                q.write_g(str(d / f"q{i:04d}.g"))
                items.append({"file": f"q{i:04d}.g", "shape": shape, "n": n, "count": counts[i], "balg": balg[i]})
            (d / "manifest.json").write_text(json.dumps({"config": 3, "seed": seed, "budget": budget, "queries": items,
                                                         "best": tune.best, "best_ms": tune.best_ms}))
    best = tune.best
    log(f"autotune: best {best} {tune.best_ms:.2f} ms over {len(tune.trials)} configurations")
    # plain: one launch per query, with the library's single-query defaults
    # (the fastest measured for one query; the batch's tuned knobs are slower
    # there)
    for q in queries[:16]:
        engine.match(g, q, timeout_ms=300000).close()
    tp = time.perf_counter()
    plain_counts = []
    for q in queries:
        r = engine.match(g, q, timeout_ms=300000)
        plain_counts.append(r.count(0))
        r.close()
    plain_s = time.perf_counter() - tp
    for _ in range(warmup):
        engine.match_batch(g, queries, timeout_ms=300000, **best).close()
    step_times, kms = [], []
    rb = None
    for _ in range(steps):
        if rb is not None:
            rb.close()
        ts = time.perf_counter()
        rb = engine.match_batch(g, queries, timeout_ms=300000, **best)
        step_times.append(time.perf_counter() - ts)
        kms.append(rb.stats().kernel_ms)
    st = rb.stats()
    ok = all(rb.query(i).count == counts[i] == plain_counts[i] and rb.query(i).status == 0 and
             rb.query(i).rows == counts[i] for i in range(len(queries)))
    embeddings = sum(rb.query(i).count for i in range(len(queries)))
    rb.close()
    ms_step = 1e3 * float(np.mean(step_times))
    value = len(queries) / (ms_step / 1e3)
    dom_ms = float(np.mean(kms))
    achieved = sum(balg) / (dom_ms / 1e3) / 1e9
    mix = {name: sum(1 for c in chosen if c[0] == k) for k, name in ((0, "tree"), (1, "cycle+chord"),
                                                                     (2, "near-clique"))}
    out = {
        "queries": len(queries), "batched_queries_per_s": value, "batched_ms_per_launch": ms_step,
        "plain_queries_per_s": len(queries) / plain_s, "batched_over_plain": value / (len(queries) / plain_s),
        "autotune": {"best": best, "best_kernel_ms": tune.best_ms, "trials": len(tune.trials)},
        "shape_mix": mix,
        "shape_mix_note": "caps 40/40/20; trees with at most the table budget of embeddings are rare on this "
                          "graph, and their share is filled by the other shapes",
        "table_budget": budget, "embeddings_per_launch": embeddings,
        "embeddings_per_s": embeddings / (ms_step / 1e3), "wave_busy_frac": st.busy_frac,
        "launch": {"tasks": int(st.tasks_pushed), "tasks_split": int(st.tasks_split), "requeues": int(st.requeues)},
        "parity_checks_passed": bool(ok),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "alg_bytes_per_launch": sum(balg), "kernel": "fm_match_kernel_batch",
                     "kernel_ms_per_launch": dom_ms},
    }
    # HBM bytes of the batched launch, measured with rocprofv3 FETCH_SIZE /
    # WRITE_SIZE on this query set and calibrated against a copy
    tpath = ROOT / "profiles" / "fm_match_kernel_rmat20_batched_traffic.json"
    if tpath.exists():
        tr = json.loads(tpath.read_text())
        if tr.get("queries") == len(queries) and tr.get("embeddings_per_launch") == embeddings:
            out["roofline"]["traffic"] = tr["hbm_bytes_per_launch"]
            out["roofline"]["traffic_source"] = str(tpath.relative_to(ROOT))
    engine.close()
    return out, chosen


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--seed", type=int, default=0xF0F2)
    ap.add_argument("--budget", type=int, default=0,
                    help="max embeddings per query with a table (default 2^26, 2^22 for config 3)")
    ap.add_argument("--tree-budget", type=int, default=1 << 31,
                    help="config 2: max embeddings of a count-only tree query")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU baseline: seconds per leg")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="config 2: leave out the batched leg")
    ap.add_argument("--queries", type=int, default=0)
    ap.add_argument("--save-queries", default="", help="write the selected queries + manifest here")
    ap.add_argument("--load-queries", default="", help="reuse a saved query set (no screening launches)")
    ap.add_argument("--rates", default="200,1000,5000", help="config 4: offered loads in requests/s")
    ap.add_argument("--serve-seconds", type=float, default=6.0, help="config 4: arrival window per rate")
    ap.add_argument("--intersect", action="store_true",
                    help="only the intersection microkernel (2,048 pairs of ~32k-element lists); one JSON line")
    args = ap.parse_args()
    if not args.budget:
        args.budget = {3: 1 << 22, 4: 1 << 18, 5: 1 << 27}.get(args.config, 1 << 26)
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on one MI355X only: there is no multi-GPU path "
              "(the data graph would be replicated and root ranges sharded)", file=sys.stderr)
        sys.exit(2)
    if fm.device_count() == 0:
        raise SystemExit("no HIP device")

    if args.intersect:
        with fm.Engine() as engine:
            res = intersect_bench(engine, reps=max(1, args.steps))
        print(json.dumps({"metric": "intersection GB/s (lists read once, intersection written once)",
                          "value": res["achieved"], "unit": "GB/s", "n_gpus": 1, "steps": max(1, args.steps),
                          "higher_is_better": True, "dtype": "u32", "data": "synthetic (seeded sorted lists)",
                          "config": {"workload": f"intersect_{res['pairs']}pairs_32k"},
                          "roofline": {"bound": "hbm", "achieved": res["achieved"], "peak": HBM_PEAK_GBS,
                                       "unit": "GB/s", "frac": res["frac"], "traffic": res["traffic"]},
                          "kernel": res}), flush=True)
        return

    t0 = time.time()
    g = build_graph(args.config, args.seed)
    g.upload()
    gi = g.info()
    log(f"graph: V={gi.vertices} E={gi.edges} labels={gi.labels} built in {time.time() - t0:.1f}s")
    graph_desc = f"rmat scale {gi.vertices.bit_length() - 1}, {gi.edges} directed edges, {gi.labels} labels"

    if args.config == 3:
        res, chosen = run_batched(g, args.seed, args.queries or 1024, args.budget, args.steps, args.warmup,
                                  save=args.save_queries, load=args.load_queries)
        if not args.load_queries:
            with fm.Engine() as engine:
                res["intersect_roofline"] = intersect_bench(engine)
        cpu = None
        if not args.no_cpu:
            cpu = cpu_baselines(g, chosen[:32], args.cpu_seconds)
        line = {"metric": "queries/s (batched, one persistent launch)", "value": res.pop("batched_queries_per_s"),
                "unit": "queries/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                "ms_per_step": res.pop("batched_ms_per_launch"), "higher_is_better": True, "scaling": "weak",
                "vs_baseline": None, "dtype": "u32", "data": "synthetic (seeded R-MAT graph and random-walk queries)",
                "config": {"workload": f"rmat20_batched_{len(chosen)}q", "graph": graph_desc,
                           "queries": len(chosen), "budget": args.budget, "seed": args.seed},
                "roofline": res.pop("roofline"), "cpu_baseline": cpu}
        line.update(res)
        print(json.dumps(line), flush=True)
        return

    if args.config == 4:
        # CurBankSize = 1 << 20 (8 MiB banks); 64 GiB of banks, 64 GiB of pages
        engine = fm.Engine(soln_prealloc=64 << 30, output_prealloc=64 << 30, bank_size=1 << 20,
                           prealloc=BENCH_PREALLOC)
    elif args.config == 5:
        # tables up to ~100 GB resident at once
        engine = fm.Engine(soln_prealloc=48 << 30, output_prealloc=160 << 30, prealloc=BENCH_PREALLOC)
    else:
        engine = fm.Engine()
    ei = engine.info()
    log(f"engine: {ei.device_name.decode()} CUs={ei.compute_units} waves/CU={ei.waves_per_cu} "
        f"grid={ei.grid_waves} banks={ei.banks}x{ei.bank_bytes}B")

    nq = args.queries or {2: 32, 4: 512, 5: 64}[args.config]
    t1 = time.time()
    manifest = None
    tables = None
    if args.load_queries:
        manifest = json.loads((Path(args.load_queries) / "manifest.json").read_text())
        chosen = [(m["shape"], m["n"], fm.Query.load_g(str(Path(args.load_queries) / m["file"])))
                  for m in manifest["queries"]]
        counts = [m["count"] for m in manifest["queries"]]
        tables = [m.get("table", True) for m in manifest["queries"]]
    elif args.config == 2:
        chosen, counts, tables = select_plain(engine, g, args.seed, nq, args.budget, args.tree_budget)
    elif args.config == 5:
        # Dense 10-vertex patterns on this graph have far more than 10^9
        # embeddings each (the hub core of the R-MAT graph holds them), so
        # none is screened: each query's table is capped instead (below).
        chosen = make_candidates(g, 5, args.seed, nq)
        counts = [None] * len(chosen)
    else:
        chosen, counts = select_mixed(engine, g, args.config, args.seed, nq, args.budget)
    if tables is None:
        tables = [True] * len(chosen)
    log(f"selected {len(chosen)} queries in {time.time() - t1:.1f}s; "
        f"total embeddings {sum(counts) if None not in counts else 'capped'}")
    queries = [c[2] for c in chosen]
    ns = [c[1] for c in chosen]

    # algorithmic bytes (instrumented, untimed)
    log("instrumented run for the algorithmic bytes")
    if args.config == 5:
        balg, rows_bytes = None, None  # capped enumeration: no complete algorithmic count
    elif manifest is not None:
        balg = [m["balg"] for m in manifest["queries"]]
        rows_bytes = sum(m["row_bytes"] for m in manifest["queries"])
    else:
        ri = engine.match_batch(g, queries, table=False, instrument=True, timeout_ms=300000)
        balg = [b_alg(ri.query(i), ns[i], table=tables[i]) for i in range(len(queries))]
        row_b = [ri.query(i).alg_row_bytes for i in range(len(queries))]
        rows_bytes = sum(row_b)
        ri.close()
        if args.save_queries:
            d = Path(args.save_queries)
            d.mkdir(parents=True, exist_ok=True)
            items = []
            for i, (shape, n, q) in enumerate(chosen):
                q.write_g(str(d / f"q{i:04d}.g"))
                items.append({"file": f"q{i:04d}.g", "shape": shape, "n": n, "count": counts[i], "balg": balg[i],
                              "row_bytes": row_b[i], "table": tables[i]})
            (d / "manifest.json").write_text(json.dumps({"config": args.config, "seed": args.seed,
                                                         "budget": args.budget, "queries": items}))

    result = {}
    if args.config == 2:
        pr = run_plain(engine, g, queries, tables, counts, args.steps, args.warmup, validate=manifest is None)
        ntree = sum(1 for t in tables if not t)
        workload = f"rmat20_plain_{len(queries)}q_{ntree}treelike_count_{len(queries) - ntree}cyclic_table"
        dom_ms = float(np.sum(pr["kernel_ms"]))
        nl = len(queries)
        achieved = sum(balg) / (dom_ms / 1e3) / 1e9
        embeddings = pr["embeddings_per_step"]
        result = {
            "metric": "embeddings/s (plain, one launch per query)",
            "value": embeddings / (pr["ms_per_step"] / 1e3),
            "unit": "embeddings/s",
            "ms_per_step": pr["ms_per_step"],
            "queries_per_s": nl / (pr["ms_per_step"] / 1e3),
            "kernel_embeddings_per_s": embeddings / (dom_ms / 1e3),
            "latency_ms_p50": pr["latency_ms_p50"],
            "latency_ms_p99": pr["latency_ms_p99"],
            "wave_busy_frac": pr["wave_busy_frac"],
            "parity_checks_passed": pr["parity_checks_passed"],
            "embeddings_per_step": embeddings,
            "count_only_treelike": pr["count_only_trees"],
            "treelike_shapes": {"tree": sum(1 for c, t in zip(chosen, tables) if not t and c[0] == 0),
                                "one_cycle": sum(1 for c, t in zip(chosen, tables) if not t and c[0] == 3)},
            "cyclic_tables": pr["cyclic_tables"],
            "launch_per_step": pr["per_step"],
            # per launch: algorithmic bytes (mean over the step's launches)
            # over the kernel's mean launch time; traffic (rocprofv3) per
            # launch too
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                         "alg_bytes_per_launch": sum(balg) / nl, "kernel_ms_per_launch": dom_ms / nl,
                         "launches_per_step": nl, "alg_bytes_per_step": sum(balg),
                         "row_bytes_per_step": rows_bytes, "kernel": "fm_match_kernel<true>",
                         "kernel_ms_per_step": dom_ms},
        }
        # HBM bytes per launch, measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
        # on this workload and calibrated against a copy of known size
        tpath = ROOT / "profiles" / "fm_match_kernel_rmat20_plain_traffic.json"
        if tpath.exists():
            tr = json.loads(tpath.read_text())
            if tr.get("workload") == workload and tr.get("embeddings_per_step") == embeddings:
                result["roofline"]["traffic"] = tr["hbm_bytes_per_launch"]
                result["roofline"]["traffic_per_step"] = tr["hbm_bytes_per_launch"] * nl
                result["roofline"]["traffic_source"] = str(tpath.relative_to(ROOT))
    elif args.config == 4:
        rng = np.random.default_rng(args.seed)
        # tables of 8 queries matched alone, for the mix check (untimed)
        alone = []
        for q in queries[:8]:
            r = engine.match(g, q, timeout_ms=300000)
            alone.append((r.count(0), r.table(0, sorted=True)))
            r.close()
        rates = [float(x) for x in args.rates.split(",") if x]
        per_rate = []
        banks = engine.info().banks
        with fm.Server(engine, g, slots=256, slot_pages=4096, bank_budget=max(1, banks * 3 // 4),
                       max_launch_ms=30000) as srv:
            serve_rate(srv, queries, counts, rates[0], 1.0, rng)  # warmup
            for rate in rates:
                pr = serve_rate(srv, queries, counts, rate, args.serve_seconds, rng, balg=balg)
                log(f"serve {rate:.0f} req/s: {pr['achieved_queries_per_s']:.1f} q/s, "
                    f"p50 {pr['latency_ms_p50']:.2f} ms, p99 {pr['latency_ms_p99']:.2f} ms, "
                    f"mismatches {pr['mismatches']}")
                per_rate.append(pr)
            mix_ok = serve_mix_check(srv, queries, alone, rng)
            sst = srv.stats()
        top = per_rate[-1]
        ok = mix_ok and all(pr["mismatches"] == 0 for pr in per_rate) and sst.dropped_tasks == 0
        achieved = top["alg_gbs_wall"]
        result = {
            "metric": f"queries/s (serving, Poisson arrivals at {top['offered_req_per_s']:.0f} req/s)",
            "value": top["achieved_queries_per_s"],
            "unit": "queries/s",
            "ms_per_step": 1e3 / top["achieved_queries_per_s"],
            "latency_ms_p50": top["latency_ms_p50"],
            "latency_ms_p99": top["latency_ms_p99"],
            "rates": per_rate,
            "server": {"slots": 256, "bank_bytes": engine.info().bank_bytes, "banks": banks,
                       "bank_budget": max(1, banks * 3 // 4), "launches": int(sst.launches),
                       "refused": int(sst.refused), "dropped_tasks": int(sst.dropped_tasks)},
            "mix_equals_alone": bool(mix_ok),
            "parity_checks_passed": bool(ok),
            "shape_mix": {name: sum(1 for c in chosen if c[0] == k)
                          for k, name in ((0, "tree"), (1, "cycle+chord"), (2, "near-clique"))},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                         "alg_bytes_per_step": top["alg_bytes"] / top["requests"],
                         "kernel": "fm_match_kernel (persistent serving launch, wall clock)",
                         "kernel_ms_per_step": 1e3 / top["achieved_queries_per_s"]},
        }
        workload = f"rmat22_serving_{len(queries)}q_poisson"
    elif args.config == 5:
        # batched only, every table resident in the page arena at once. Each
        # query stops once it has more than `cap` embeddings (status
        # COUNT_LIMIT) with `cap` rows stored: 64 x 3.9e7 rows of 10 u32 =
        # 100 GB of tables per step.
        cap = int(2.5e9) // len(queries)
        kw = dict(max_rows=cap, max_count=cap, timeout_ms=300000, allow_errors=True)
        for _ in range(args.warmup):
            engine.match_batch(g, queries, **kw).close()
        step_times, kms, rb = [], [], None
        for _ in range(args.steps):
            if rb is not None:
                rb.close()
            ts = time.perf_counter()
            rb = engine.match_batch(g, queries, **kw)
            step_times.append(time.perf_counter() - ts)
            kms.append(rb.stats().kernel_ms)
            log(f"stress step: {step_times[-1] * 1e3:.0f} ms")
        st = rb.stats()
        outs = [rb.query(i) for i in range(len(queries))]
        # a query either completed (exact count, every row stored) or passed
        # the cap (its table full); every sampled stored row is an embedding
        ok = st.status == 0 and all(
            (o.status == 0 and o.rows == o.count) or (o.status == FM_ERR_COUNT_LIMIT and o.rows == cap)
            for o in outs)
        rng = np.random.default_rng(5)
        for i in range(0, len(queries), 4):
            ok = ok and validate_sample(g, queries[i], rb.table(i, sorted=False, max_rows=1 << 16), rng)
        rows = sum(o.rows for o in outs)
        ms_step = 1e3 * float(np.mean(step_times))
        dom_ms = float(np.mean(kms))
        table_bytes = int(sum(4 * 10 * o.rows for o in outs))
        achieved = table_bytes / (dom_ms / 1e3) / 1e9
        result = {
            "metric": "embeddings/s (batched stress, tables resident; rows stored)",
            "value": rows / (ms_step / 1e3),
            "unit": "embeddings/s",
            "ms_per_step": ms_step,
            "queries_per_s": len(queries) / (ms_step / 1e3),
            "rows_per_step": rows,
            "row_cap_per_query": cap,
            "queries_at_cap": sum(1 for o in outs if o.status != 0),
            "table_bytes_per_step": table_bytes,
            "wave_busy_frac": st.busy_frac,
            "launch": {"tasks": int(st.tasks_pushed), "banks_fresh": int(st.banks_allocated),
                       "banks_reused": int(st.banks_reused), "bank_alloc_failures": int(st.alloc_failures),
                       "requeues": int(st.requeues), "splits": int(st.tasks_split),
                       "pages": int(st.pages_allocated)},
            "parity_checks_passed": bool(ok),
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                         "alg_bytes_per_step": table_bytes, "alg_bytes_note": "table bytes only (capped run)",
                         "kernel": "fm_match_kernel", "kernel_ms_per_step": dom_ms},
        }
        workload = f"rmat24_stress_{len(queries)}q_dense10"
        # Where the step's work goes (untimed): the same capped launch counted
        # without tables and with the engine's counters, summed over the
        # queries: partial embeddings per depth (depth 9 = embeddings found
        # before the cap stopped each query), candidates examined, membership
        # probes and neighbour-row words read.
        rb.close()
        rb = None
        ri = engine.match_batch(g, queries, table=False, instrument=True, max_count=cap, timeout_ms=300000,
                                allow_errors=True)
        qo = [ri.query(i) for i in range(len(queries))]
        depth = max(len(o.partials) for o in qo)
        parts = [sum(o.partials[d] for o in qo if d < len(o.partials)) for d in range(depth)]
        result["work_breakdown"] = {
            "partials_per_depth": parts,
            "candidates": sum(o.candidates for o in qo), "probes": sum(o.probes for o in qo),
            "row_words": sum(o.alg_row_bytes for o in qo) // 4,
            "kernel_ms": ri.stats().kernel_ms,
            "note": "count mode, counters on, same cap; partials at depth d are the valid partial embeddings "
                    "of d + 1 vertices, the last entry the embeddings found"}
        ri.close()
        log(f"stress breakdown: partials per depth {parts}")

    if not args.load_queries:
        result["intersect_roofline"] = intersect_bench(engine)
    engine.close()

    if args.config == 2 and not args.no_batched and not args.load_queries:
        # the batched form on the same graph (its own engine: 1,024 tables)
        res, _ = run_batched(g, args.seed, 1024, 1 << 22, args.steps, args.warmup)
        result["batched"] = res

    cpu = None
    # config 5: the oracle's sorted copy of 2^29 edge lines alone takes minutes
    # of host time, so the stress line carries no CPU baseline
    if not args.no_cpu and args.config != 5:
        cpu = cpu_baselines(g, chosen[: min(len(chosen), 32)], args.cpu_seconds)

    line = {
        "metric": result.pop("metric"),
        "value": result.pop("value"),
        "unit": result.pop("unit"),
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": result.pop("ms_per_step"),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded R-MAT graph and random-walk queries)",
        "config": {"workload": workload, "graph": graph_desc, "queries": len(queries), "budget": args.budget,
                   "seed": args.seed},
        "roofline": result.pop("roofline"),
        "cpu_baseline": cpu,
    }
    if args.config == 2:
        line["config"]["tree_budget"] = args.tree_budget
    line.update(result)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
