"""Benchmark of the MI355X InsPIRe respond path (one GPU).

Headline workload: the batched exact-gadget TwoPacking respond over a 2^20 x 512-byte database
(512 MiB, 512 shards of 2048 records, gamma = 256 columns), d = 2048, q = 2^60-2^14+1, every
shard's packing store precomputed on the device, one seated client key, batches of 64 seeded
queries at independent uniform record indices, 8 batches merged into each launch. A step is one
launch (512 queries answered end to end: DB scan + packing recursion + online gather, 36-bit
modulus switch); inputs are resident in HBM when the timed region starts.

Beside it the same run measures the plain path (one query per launch) over the same database,
the 2^18-record batched config (64 and 512 queries per launch), the 2^16-record plain config,
the served configs (approximate gadget over 2^20 records; special modulus), and the CPU oracle
(a C restatement of the reference algorithm) on bounded samples, and reports per-kernel
roofline figures from HIP events.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "inspire-rs_amd"))

N = 2048
GAMMA = 256
ENTRY = 512
Q36 = 68718428161
HBM_PEAK_GBS = 8000.0
VALU_MAD_PEAK_T = 256 * 64 * 2.4e9 / 1e12  # 39.3 T v_mad_u64_u32 lane-ops/s
# VALU instruction issue: one wave instruction (64 lanes on a 16-lane SIMD) per 4 cycles per SIMD
VALU_ISSUE_PEAK = 256 * 4 * 2.4e9 / 4  # 614.4 G wave instructions/s


def valu_issue(insts_per_launch, kernel_ms, counter_source):
    """Fraction of peak VALU instruction issue of a kernel: its VALU wave instructions per launch
    (rocprofv3 SQ_INSTS_VALU of the same workload, committed beside the profile; the count is fixed
    by the launch's shape) over its measured kernel time."""
    if insts_per_launch is None:
        return None
    rate = insts_per_launch / (kernel_ms * 1e-3)
    return {"achieved": round(rate / 1e9, 2), "peak": round(VALU_ISSUE_PEAK / 1e9, 1), "unit": "G wave instructions/s",
            "frac": round(rate / VALU_ISSUE_PEAK, 4), "insts_per_launch": round(insts_per_launch),
            "source": counter_source}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_db(g, ctx, num_shards, seed):
    rng = np.random.default_rng(seed)
    for s in range(num_shards):
        recs = rng.integers(0, 256, N * ENTRY, dtype=np.uint8).tobytes()
        ctx.upload_records(s, recs, ENTRY).precompute_packing()


def make_batches(rng, count, batch, num_shards, q):
    ids = rng.integers(0, num_shards, (count, batch), dtype=np.uint32)
    seeds = rng.integers(0, 256, (count, batch, 32), dtype=np.uint8)
    # b = -a*s + e + delta*m is uniform-looking mod q; draw it uniform
    b = rng.integers(0, q, (count, batch, N), dtype=np.uint64)
    return ids, seeds, b


def _cpu_threads():
    """Host threads for the all-core CPU legs: the box's share (at most 16 per GPU)."""
    return max(1, min(os.cpu_count() or 1, 16))


def _timed_queries(run_one, budget_s, threads):
    """Queries per second of run_one(worker_rng) on `threads` host threads for about budget_s
    (the oracle's C calls release the GIL, so the threads run in parallel)."""
    import threading

    counts = [0] * threads
    t0 = time.perf_counter()

    def worker(k):
        rng = np.random.default_rng(1000 + k)
        while time.perf_counter() - t0 < budget_s or counts[k] == 0:
            run_one(rng)
            counts[k] += 1

    ts = [threading.Thread(target=worker, args=(k,)) for k in range(threads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    dt = time.perf_counter() - t0
    return sum(counts), dt


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def _pct(xs):
    xs = np.sort(np.asarray(xs, dtype=np.float64))
    return {k: round(float(np.percentile(xs, q)), 3) for k, q in (("p50", 50), ("p95", 95), ("p99", 99))}


def cpu_baseline(budget_s=10.0, rounds=30):
    """The CPU oracle (C restatement of the reference respond) on bounded samples.

    value: seeded d2048 gamma=256 exact-gadget queries against a prebuilt store (the headline's
    per-query work) on all of the box's host cores, one query per thread as a multi-threaded
    server answers independent requests, with the reference's AVX-512 IFMA multiply-accumulates
    (its simd-packing-offline feature) where the host has them; scalar and one-core rates beside.

    latency: the reference benches' method (respond_latency_bench.rs:29-86): warm-up, then
    `rounds` rounds, each timing one query of every config on one thread in an order that
    rotates from round to round, p50/p95/p99 per config and the 1-minute load average at each
    round:
      C0  d2048 gamma 16, 2048 x 32 B: respond_seeded_inspiring uncached (PackParams and the
          offline keys built per call, no store) then extract_inspiring;
      C2  d2048 gamma 256 exact gadget against a prebuilt store (C1-C2 and the headline);
      C3  d2048_approximate_gadget gamma 256 against a prebuilt store (the served config);
      C4  special modulus ring2048_cell256 against one prebuilt table and seat.
    The oracle answers a query on one thread (the reference's rayon sections, the column fold
    and step 1, are a few percent of a query; its recursion is serial)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    rng = np.random.default_rng(99)
    scalar = oracle.Pack(N, GAMMA)
    simd = oracle.Pack(N, GAMMA)
    has_simd = simd.set_simd(True)
    cols = rng.integers(0, 1 << 16, (GAMMA, N), dtype=np.uint64)
    store = scalar.store_build(cols)
    y = rng.integers(0, oracle.Q, (3, N), dtype=np.uint64)
    w_seed = bytes(32)

    def runner(p):
        def one(r):
            seed = r.integers(0, 256, 32, dtype=np.uint8).tobytes()
            b = r.integers(0, oracle.Q, N, dtype=np.uint64)
            p.respond(w_seed, oracle.expand_seed(seed, N), b, cols, y, store)
        return one

    T = _cpu_threads()
    legs = [("scalar", scalar)] + ([("simd", simd)] if has_simd else [])
    share = budget_s / (2 * len(legs))
    res = {}
    for name, p in legs:
        n1, dt1 = _timed_queries(runner(p), share, 1)
        nT, dtT = _timed_queries(runner(p), share, T)
        res[name] = (n1, dt1, nT, dtT)
    best = "simd" if has_simd else "scalar"
    n1, dt1, nT, dtT = res[best]
    sn1, sdt1, snT, sdtT = res["scalar"]
    form = "with AVX-512 IFMA multiply-accumulates" if has_simd else "scalar (host has no AVX-512 IFMA)"

    # --- per-query latency of each config, rounds in rotating order
    c0 = oracle.Cell(N, 32, 0x5EED_C0)
    c0_queries = [c0.query(int(i)) for i in np.random.default_rng(5).integers(0, 3 * N, rounds + 2)]
    approx = oracle.Pack(N, GAMMA, 3, 1)
    approx.set_simd(has_simd)
    y2 = rng.integers(0, oracle.Q, (2, N), dtype=np.uint64)
    astore = approx.store_build(cols)
    ring = oracle.SmRing(cell_bytes=SM_CELL, record_bytes=SM_RECORD)
    table = ring.table(sm_cells(rng, SM_N))
    sm_seat = ring.seat(rng.integers(0, oracle.SM_Q, SM_N, dtype=np.uint32),
                        rng.integers(0, oracle.SM_P, SM_N, dtype=np.uint32), bytes(32))
    sm_avx = oracle.sm_avx512_available()
    fast = simd if has_simd else scalar

    def c0_one(k):
        shard, seed, b = c0_queries[k]
        p = oracle.Pack(N, c0.gamma)  # uncached: PackParams + offline keys per call
        a_hat, bo = p.respond(c0.w_seed, oracle.expand_seed(seed, N), b, c0.columns(shard), c0.y_body)
        return oracle.extract_inspiring(c0.secret, a_hat, bo[:c0.gamma], 32)

    def c2_one(k):
        r = np.random.default_rng(700 + k)
        fast.respond(w_seed, oracle.expand_seed(r.bytes(32), N), r.integers(0, oracle.Q, N, dtype=np.uint64), cols,
                     y, store)

    def c3_one(k):
        r = np.random.default_rng(800 + k)
        approx.respond(w_seed, oracle.expand_seed(r.bytes(32), N), r.integers(0, oracle.Q, N, dtype=np.uint64),
                       cols, y2, astore)

    def c4_one(k):
        r = np.random.default_rng(900 + k)
        ring.respond(table, sm_seat, r.bytes(32), r.integers(0, oracle.SM_Q, SM_N, dtype=np.uint32), avx512=sm_avx)

    configs = [("C0", c0_one), ("C2", c2_one), ("C3", c3_one), ("C4", c4_one)]
    for _, fn in configs:  # warm-up
        fn(rounds + 1)
    lat = {name: [] for name, _ in configs}
    loads = []
    for r in range(rounds):
        order = configs[r % len(configs):] + configs[:r % len(configs)]
        for name, fn in order:
            t0 = time.perf_counter()
            fn(r)
            lat[name].append((time.perf_counter() - t0) * 1e3)
        loads.append(round(os.getloadavg()[0], 2))
    latency = {name: {"ms": _pct(v), "queries_per_s_one_core": round(1e3 / float(np.mean(v)), 2)}
               for name, v in lat.items()}
    latency["C0"]["workload"] = ("d2048 gamma 16, 2048 x 32 B: respond_seeded_inspiring uncached (PackParams and "
                                 "offline keys built per call, per-query packing from the columns) + "
                                 "extract_inspiring")
    latency["C2"]["workload"] = "d2048 gamma 256 exact gadget, prebuilt store" + (" (IFMA)" if has_simd else "")
    latency["C3"]["workload"] = "d2048_approximate_gadget gamma 256, prebuilt store" + (" (IFMA)" if has_simd else "")
    latency["C4"]["workload"] = ("special modulus ring2048_cell256, prebuilt table and seat"
                                 + (" (AVX-512 build)" if sm_avx else ""))
    return {"value": round(nT / dtT, 3), "unit": "queries/s", "cores": T, "kind": "port",
            "cpu_model": _cpu_model(), "single_core": round(n1 / dt1, 3),
            "simd": "avx512ifma" if has_simd else None,
            "without_simd": {"value": round(snT / sdtT, 3), "single_core": round(sn1 / sdt1, 3)},
            "sample": f"seeded d2048 gamma=256 queries against a prebuilt store (C restatement), {form}: "
                      f"{nT} on {T} threads in {dtT:.1f} s and {n1} on 1 thread in {dt1:.1f} s; scalar "
                      f"{snT} on {T} threads in {sdtT:.1f} s and {sn1} on 1 thread in {sdt1:.1f} s",
            "latency": {"rounds": rounds, "order": "rotating per round over C0, C2, C3, C4", "threads": 1,
                        "load_avg_1min": {"min": min(loads), "max": max(loads), "per_round": loads},
                        **latency}}


def served(g, seconds, rate_qps, seed=7, capacity_queries=0, queries_per_cu=2):
    """Served config: d2048_approximate_gadget, 2^20 x 512 B records (512 shards, stores
    precomputed), 1024 seated client keys registered by handle, Poisson arrivals of
    batches of 1-32 queries at uniform indices answered by the continuous-batching
    server. Inputs come from host memory (PCIe-inclusive); latency is submit to result.

    Three runs over the same server: arrivals paced at rate_qps (the latency operating
    point); the same pacing with about 10% of requests opening a new session first (the
    key is seated by register_seat while launches run, and the request's latency counts
    the registration); then, when capacity_queries > 0, that many queries submitted at
    once as a backlog, whose drain rate is the server's capacity."""
    import threading

    rng = np.random.default_rng(seed)
    w_seed = rng.integers(0, 256, 32, dtype=np.uint8).tobytes()
    ctx = g.Context(g.params_d2048(approximate=True), w_seed, GAMMA)
    ctx.set_queries_per_cu(queries_per_cu)
    shards = 512
    t0 = time.perf_counter()
    for s in range(shards):
        ctx.upload_records(s, rng.integers(0, 256, N * ENTRY, dtype=np.uint8).tobytes(), ENTRY).precompute_packing()
    build = time.perf_counter() - t0
    server = g.Server(ctx, max_launch=512)
    handles = [server.register_seat(rng.integers(0, g.Q, (2, N), dtype=np.uint64)) for _ in range(1024)]
    # query bodies come from a pool: their values do not change the work
    pool = rng.integers(0, g.Q, (64, 32, N), dtype=np.uint64)

    # key rows of the sessions opened during the registering run (seating copies them)
    new_keys = rng.integers(0, g.Q, (8, 2, N), dtype=np.uint64)

    def schedule(total_s=None, total_q=None, new_frac=0.0):
        reqs, t, q = [], 0.0, 0
        while (total_s is not None and t < total_s) or (total_q is not None and q < total_q):
            count = int(rng.integers(1, 33))
            t += rng.exponential(count / rate_qps)
            q += count
            reqs.append((t, int(rng.integers(0, len(handles))), rng.integers(0, shards, count).astype(np.uint32),
                         [rng.integers(0, 256, 32, dtype=np.uint8).tobytes() for _ in range(count)],
                         pool[int(rng.integers(0, len(pool)))][:count], bool(rng.random() < new_frac)))
        return reqs

    def run(reqs, paced):
        lats, done_q = [], [0]
        new_lats, reg_ms = [], []
        lock = threading.Lock()
        tickets = []

        def waiter():
            while True:
                with lock:
                    if not tickets:
                        if waiter.stop:
                            return
                        tk = None
                    else:
                        tk = tickets.pop(0)
                if tk is None:
                    time.sleep(0.0005)
                    continue
                tk, reg = tk
                a, b, lat = server.wait(tk)
                with lock:
                    lats.append(lat + reg)
                    if reg > 0:
                        new_lats.append(lat + reg)
                    done_q[0] += a.shape[0]

        waiter.stop = False
        ws_threads = [threading.Thread(target=waiter) for _ in range(4)]
        for th in ws_threads:
            th.start()
        launches0 = server.stats()["launches"]
        start = time.perf_counter()
        for (at, h, ids, seeds, bs, new) in reqs:
            if paced:
                delay = start + at - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
            reg = 0.0
            if new:
                # a new session: its key is seated before its first request is submitted
                t_reg = time.perf_counter()
                handle = server.register_seat(new_keys[len(reg_ms) % len(new_keys)])
                reg = (time.perf_counter() - t_reg) * 1e3
                reg_ms.append(reg)
            else:
                handle = handles[h]
            tk = server.submit(handle, ids, seeds, bs, switch_modulus=Q36)
            with lock:
                tickets.append((tk, reg))
        waiter.stop = True
        for th in ws_threads:
            th.join()
        elapsed = time.perf_counter() - start
        run.new_lats, run.reg_ms = np.array(new_lats), np.array(reg_ms)
        return done_q[0], elapsed, np.array(lats), server.stats()["launches"] - launches0

    def pct(x):
        return {"p50": round(float(np.percentile(x, 50)), 3), "p99": round(float(np.percentile(x, 99)), 3)}

    reqs = schedule(total_s=seconds)
    # warm the server
    server.wait(server.submit(handles[0], reqs[0][2], reqs[0][3], reqs[0][4], switch_modulus=Q36))
    done, elapsed, lats, launches = run(reqs, paced=True)
    seats_before = server.stats()["seats"]
    rreqs = schedule(total_s=seconds, new_frac=0.1)
    rdone, relapsed, rlats, rlaunches = run(rreqs, paced=True)
    registering = {"value": round(rdone / relapsed, 1), "unit": "queries/s", "requests": len(rlats),
                   "new_sessions": len(run.reg_ms), "latency_ms": pct(rlats),
                   "new_session_latency_ms": pct(run.new_lats) if len(run.new_lats) else None,
                   "register_ms": pct(run.reg_ms) if len(run.reg_ms) else None,
                   "seats": {"before": seats_before, "after": server.stats()["seats"]},
                   "workload": "the paced mix with about 10% of requests opening a new session first: its key "
                               "is seated (register_seat) while launches run, and the request's latency "
                               "includes the registration"}
    capacity = None
    if capacity_queries > 0:
        creqs = schedule(total_q=capacity_queries)
        cdone, celapsed, clats, claunches = run(creqs, paced=False)
        capacity = {"value": round(cdone / celapsed, 1), "unit": "queries/s", "queries": cdone,
                    "requests": len(creqs), "launches": claunches, "seconds": round(celapsed, 3),
                    "mean_queries_per_launch": round(cdone / max(claunches, 1), 1),
                    "workload": "the same request mix submitted at once as a backlog; value = drain rate"}
    stats = server.stats()
    server.close()
    ctx.close()
    return {"offered_qps": round(rate_qps, 1), "value": round(done / elapsed, 1), "unit": "queries/s",
            "latency_ms": pct(lats),
            "requests": len(lats), "launches": launches,
            "seats": stats["seats"], "seats_resident_bytes": stats["resident_bytes"],
            "seats_payload_bytes_reference_accounting": stats["payload_bytes"],
            "store_build_s": round(build, 2),
            "registering": registering,
            "capacity": capacity,
            "parity": "approximate-gadget response bytes: no reference vector exists for them; they rest on the "
                      "rounding contract of the device digit split, the mask-stream separation, approximate != "
                      "exact and store == per-query on the GPU, decode correctness of device responses, and "
                      "byte equality with the oracle's restatement",
            "workload": "d2048_approximate_gadget, 2^20 x 512 B (512 shards), 1024 seated keys, Poisson arrivals "
                        "of 1-32-query batches at 70% of the batched rate, host inputs (PCIe-inclusive), "
                        "continuous batching <= 512 queries per launch, 36-bit switch"}


SM_N, SM_GAMMA, SM_CELL, SM_RECORD = 2048, 512, 256, 246
# 32-bit modular products per special-modulus query: 511 steps x (4 transforms of 11264
# butterflies + 6 x 2048 pointwise), plus NTT(a), the first step and the final 3 transforms
SM_MM_PER_QUERY = 511 * (4 * 11264 + 6 * 2048) + 11264 + 2048 + 3 * 11264 + 492
# v_mul_lo_u32 / v_mul_hi_u32 issue at 64 lanes per clock per CU (half the simple-op rate)
MUL32_PEAK_T = 256 * 64 * 2.4e9 / 1e12


def sm_cells(rng, rows):
    c = np.frombuffer(rng.bytes(rows * SM_CELL), dtype=np.uint8).reshape(rows, SM_CELL).copy()
    c[:, SM_RECORD:] = 0
    return c.tobytes()


def sm_cpu_baseline(budget_s):
    """The CPU oracle's special-modulus respond (C restatement of respond_impl) on a bounded
    sample of served-set queries against one prebuilt table and seat, on all of the box's
    host cores (one query per thread) and on one; compiled for AVX-512F/BW/DQ/VL (the
    reference's Kernel::Avx512) where the host has them, and portable. value is the all-core
    rate of the faster form."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    rng = np.random.default_rng(98)
    ring = oracle.SmRing(cell_bytes=SM_CELL, record_bytes=SM_RECORD)
    table = ring.table(sm_cells(rng, SM_N))
    seat = ring.seat(rng.integers(0, oracle.SM_Q, SM_N, dtype=np.uint32),
                     rng.integers(0, oracle.SM_P, SM_N, dtype=np.uint32), bytes(32))

    def runner(avx512):
        def one(r):
            ring.respond(table, seat, r.bytes(32), r.integers(0, oracle.SM_Q, SM_N, dtype=np.uint32), avx512=avx512)
        return one

    T = _cpu_threads()
    has_avx512 = oracle.sm_avx512_available()
    legs = [("portable", False)] + ([("avx512", True)] if has_avx512 else [])
    share = budget_s / (2 * len(legs))
    res = {}
    for name, flag in legs:
        n1, dt1 = _timed_queries(runner(flag), share, 1)
        nT, dtT = _timed_queries(runner(flag), share, T)
        res[name] = (n1, dt1, nT, dtT)
    best = max(res, key=lambda k: res[k][2] / res[k][3])
    n1, dt1, nT, dtT = res[best]
    pn1, pdt1, pnT, pdtT = res["portable"]
    out = {"value": round(nT / dtT, 3), "unit": "queries/s", "cores": T, "kind": "port",
           "single_core": round(n1 / dt1, 3), "simd": "avx512" if best == "avx512" else None,
           "without_simd": {"value": round(pnT / pdtT, 3), "single_core": round(pn1 / pdt1, 3)},
           "sample": f"ring2048_cell256 queries against one prebuilt table (C restatement), "
                     f"{'compiled for AVX-512F/BW/DQ/VL' if best == 'avx512' else 'portable'}: "
                     f"{nT} on {T} threads in {dtT:.1f} s and {n1} on 1 thread in {dt1:.1f} s; portable "
                     f"{pnT} on {T} threads in {pdtT:.1f} s and {pn1} on 1 thread in {pdt1:.1f} s"}
    if has_avx512:
        an1, adt1, anT, adtT = res["avx512"]
        out["avx512"] = {"value": round(anT / adtT, 3), "single_core": round(an1 / adt1, 3)}
    return out


def sm_two_streams(torch, srv, picks, d_seeds, d_bodies, d_mask, d_body, L, K, wall, qps, lat):
    """The special-modulus launches alternating between two streams with their own outputs (and
    job lists): a launch's first workgroups run while the previous launch's last ones drain.
    Returns the faster pipeline's (queries/s, launch latencies, name) and the two-stream figures."""
    last_mask = d_mask.clone()
    st2 = [torch.cuda.Stream(), torch.cuda.Stream()]
    outs2 = [(d_mask, d_body), (torch.empty_like(d_mask), torch.empty_like(d_body))]

    def issue2(i):
        k = i % 2
        return srv.respond_batch_device(picks[i][0], picks[i][1], d_seeds[i].data_ptr(), d_bodies[i].data_ptr(),
                                        outs2[k][0].data_ptr(), outs2[k][1].data_ptr(), stream=st2[k].cuda_stream)

    issue2(0)
    torch.cuda.synchronize()
    s2 = [torch.cuda.Event(enable_timing=True) for _ in range(K)]
    e2 = [torch.cuda.Event(enable_timing=True) for _ in range(K)]
    t0 = time.perf_counter()
    for k in range(K):
        s2[k].record(st2[(1 + k) % 2])  # the stream issue2(1 + k) launches on
        issue2(1 + k)
        e2[k].record(st2[(1 + k) % 2])
    torch.cuda.synchronize()
    wall2 = time.perf_counter() - t0
    assert torch.equal(outs2[K % 2][0], last_mask), "two-stream launch differs from the stream launch"
    lat2 = np.array([a_.elapsed_time(b_) for a_, b_ in zip(s2, e2)])
    two_streams = {"value": round(L * K / wall2, 1), "unit": "queries/s",
                   "launch_ms": {"p50": round(float(np.percentile(lat2, 50)), 3),
                                 "p99": round(float(np.percentile(lat2, 99)), 3),
                                 "unit": "one launch on its stream, overlapping the other stream's"}}
    pipeline = "stream"
    if wall2 < wall:
        pipeline, qps, lat = "two streams", L * K / wall2, lat2
    return qps, lat, pipeline, two_streams


def special_modulus(g, torch, dev, stream, args):
    """Served special-modulus set (ring2048_cell256): 2^22 cells of 256 B in 2048 step-1 tables
    (one per shard of 2048 rows), 256 seated keys with their own mask seeds, batches of queries
    at uniform shards under uniform keys, one launch per batch, inputs resident in HBM."""
    rng = np.random.default_rng(4242)
    params = g.ring2048_cell256()
    srv = g.SmServer(params)
    t0 = time.perf_counter()
    tables = [srv.table(sm_cells(rng, SM_N)) for _ in range(args.sm_tables)]
    torch.cuda.synchronize()
    table_s = time.perf_counter() - t0
    seats = [srv.seat(g.PackingKey(2, rng.bytes(32), rng.integers(0, g.SM_Q, SM_N, dtype=np.uint32),
                                   rng.integers(0, g.SM_P, SM_N, dtype=np.uint32))) for _ in range(args.sm_keys)]
    L, K = args.sm_batch, args.sm_steps
    total = 1 + K
    d_seeds = torch.from_numpy(rng.integers(0, 256, (total, L, 32), dtype=np.uint8)).to(dev)
    d_bodies = torch.from_numpy(rng.integers(0, g.SM_Q, (total, L, SM_N), dtype=np.int64).astype(np.int32)).to(dev)
    d_mask = torch.empty((L, SM_N), dtype=torch.int32, device=dev)
    d_body = torch.empty((L, 2 * SM_RECORD), dtype=torch.int32, device=dev)
    picks = [([tables[i] for i in rng.integers(0, len(tables), L)], [seats[i] for i in rng.integers(0, len(seats), L)])
             for _ in range(total)]

    def issue(i, timed=False):
        return srv.respond_batch_device(picks[i][0], picks[i][1], d_seeds[i].data_ptr(), d_bodies[i].data_ptr(),
                                        d_mask.data_ptr(), d_body.data_ptr(), stream=stream, timed=timed)

    issue(0)
    torch.cuda.synchronize()
    # the job list of each launch is uploaded by the call when its handles change; measure the
    # launches back to back on one stream with their own job lists
    # and HIP events around each launch's kernel on its stream (the trace waits for nothing)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * K)]
    srv.trace(K)
    t0 = time.perf_counter()
    for k in range(K):
        ev[2 * k].record()
        issue(1 + k)
        ev[2 * k + 1].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    lat = np.array([ev[2 * k].elapsed_time(ev[2 * k + 1]) for k in range(K)])
    kern = srv.kernel_times()
    assert len(kern) == K, "a timed launch was not traced"
    kms = float(np.mean(kern))
    qps = L * K / wall
    one_stream = {"value": round(qps, 1), "unit": "queries/s",
                  "launch_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                                "p99": round(float(np.percentile(lat, 99)), 3)}}
    two_streams, pipeline = None, "stream"
    if not args.no_two_streams:
        qps, lat, pipeline, two_streams = sm_two_streams(torch, srv, picks, d_seeds, d_bodies, d_mask, d_body, L, K,
                                                         wall, qps, lat)
    mm_s = L * SM_MM_PER_QUERY / (kms * 1e-3)
    # per query the kernel streams its table (gamma x n words) and reads its cells, seat and body
    bytes_per_query = SM_GAMMA * SM_N * 4 + SM_N * SM_CELL + 4 * SM_N * 4 + SM_N * 4 + SM_N * 4 + 2 * SM_RECORD * 4
    # HBM bytes per launch measured by rocprofv3 counters for this workload (2560 queries per
    # launch), committed beside the profile; scaled to this run's launch size
    traffic = insts = None
    tpath = os.path.join(ROOT, "profiles", "sm_respond_batched2560_ring2048_cell256_traffic.json")
    if os.path.exists(tpath):
        tj = json.load(open(tpath))
        traffic = round(tj["hbm_bytes_per_dispatch"]["inspire::k_sm_respond<11>/2560"] * L / 2560)
        insts = tj.get("valu_insts_per_dispatch", {}).get("inspire::k_sm_respond<11>/2560")
        insts = None if insts is None else insts * L / 2560
    out = {
        "value": round(qps, 1), "unit": "queries/s",
        "workload": f"special_modulus ring2048_cell256: {len(tables)} step-1 tables ({len(tables) * SM_N} cells of "
                    f"{SM_CELL} B), {len(seats)} seated keys, {L} queries per launch at uniform tables and keys, "
                    f"{K} launches back to back ({pipeline}), inputs resident in HBM",
        "queries_per_launch": L,
        "pipeline": pipeline,
        "stream_launches": one_stream,
        "two_stream_launches": two_streams,
        "launch_ms": {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3)},
        "kernel_ms": round(kms, 3),
        "table_build_s": round(table_s, 2),
        "tables_resident_bytes": sum(t.resident_bytes() for t in tables),
        "seat_bytes": seats[0].bytes(),
        "roofline": {"bound": "valu", "kernel": "k_sm_respond",
                     "achieved": round(mm_s * 3 / 1e12, 3), "peak": MUL32_PEAK_T, "unit": "T mul32/s",
                     "frac": round(mm_s * 3 / 1e12 / MUL32_PEAK_T, 4), "traffic": traffic,
                     "algorithmic_bytes": L * bytes_per_query,
                     "launch_ms": round(kms, 4), "mm32_per_query": SM_MM_PER_QUERY,
                     "timing": "HIP events on the launch stream around this kernel in each launch of the timed region",
                     "hbm_gbs": round(L * bytes_per_query / (kms * 1e-3) / 1e9, 1),
                     "definition": "3 v_mul_lo/hi_u32 per 32-bit modular product (Shoup or Montgomery), "
                                   "algorithmic; peak = 256 CU x 64 lanes x 2.4 GHz",
                     "valu_issue": valu_issue(insts, kms, "profiles/sm_respond_batched2560_ring2048_cell256.md")},
    }
    if not args.no_served:
        out["served"] = sm_served(g, srv, tables, rng, rate_qps=0.7 * qps, seconds=args.served_seconds,
                                  capacity_queries=int(max(args.served_seconds, 1.0) * qps), max_launch=L)
    for t in tables:
        t.close()
    for st in seats:
        st.close()
    srv.close()
    return out


def sm_served(g, srv, tables, rng, rate_qps, seconds, keys=256, capacity_queries=0, max_launch=2560):
    """Served special modulus: 256 keys seated by handle in a continuous-batching queue,
    Poisson arrivals of batches of 1-64 queries at uniform tables under uniform keys, from
    host memory (PCIe-inclusive); latency is submit to result. As for the gadget-key served
    leg, arrivals paced at rate_qps, then capacity_queries submitted at once as a backlog
    whose drain rate is the batcher's capacity."""
    import threading

    batcher = g.SmBatcher(srv, max_launch=max_launch)
    handles = [batcher.register_key(g.PackingKey(2, rng.bytes(32), rng.integers(0, g.SM_Q, SM_N, dtype=np.uint32),
                                                 rng.integers(0, g.SM_P, SM_N, dtype=np.uint32)))
               for _ in range(keys)]
    # query bodies come from a pool: their values do not change the work
    pool = [g.Query(rng.bytes(32), rng.integers(0, g.SM_Q, SM_N, dtype=np.uint32)) for _ in range(256)]

    def schedule(total_s=None, total_q=None):
        reqs, t, q = [], 0.0, 0
        while (total_s is not None and t < total_s) or (total_q is not None and q < total_q):
            count = int(rng.integers(1, 65))
            t += rng.exponential(count / rate_qps)
            q += count
            reqs.append((t, handles[int(rng.integers(0, keys))],
                         [tables[i] for i in rng.integers(0, len(tables), count)],
                         [pool[i] for i in rng.integers(0, len(pool), count)]))
        return reqs

    def run(reqs, paced):
        lats, done_q, tickets = [], [0], []
        lock = threading.Lock()

        def waiter():
            while True:
                with lock:
                    tk = tickets.pop(0) if tickets else None
                    if tk is None and waiter.stop:
                        return
                if tk is None:
                    time.sleep(0.0005)
                    continue
                m, b, lat = batcher.wait(tk)
                with lock:
                    lats.append(lat)
                    done_q[0] += m.shape[0]

        waiter.stop = False
        ths = [threading.Thread(target=waiter) for _ in range(4)]
        for th in ths:
            th.start()
        launches0 = batcher.stats()["launches"]
        start = time.perf_counter()
        for (at, h, tl, ql) in reqs:
            if paced:
                delay = start + at - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
            tk = batcher.submit(h, tl, ql)
            with lock:
                tickets.append(tk)
        waiter.stop = True
        for th in ths:
            th.join()
        return done_q[0], time.perf_counter() - start, np.array(lats), batcher.stats()["launches"] - launches0

    reqs = schedule(total_s=seconds)
    batcher.wait(batcher.submit(reqs[0][1], reqs[0][2], reqs[0][3]))  # warm
    done, elapsed, lats, launches = run(reqs, paced=True)
    capacity = None
    if capacity_queries > 0:
        creqs = schedule(total_q=capacity_queries)
        cdone, celapsed, _, claunches = run(creqs, paced=False)
        capacity = {"value": round(cdone / celapsed, 1), "unit": "queries/s", "queries": cdone,
                    "requests": len(creqs), "launches": claunches, "seconds": round(celapsed, 3),
                    "mean_queries_per_launch": round(cdone / max(claunches, 1), 1),
                    "workload": "the same request mix submitted at once as a backlog; value = drain rate"}
    stats = batcher.stats()
    batcher.close()
    return {"offered_qps": round(rate_qps, 1), "value": round(done / elapsed, 1), "unit": "queries/s",
            "latency_ms": {"p50": round(float(np.percentile(lats, 50)), 3),
                           "p99": round(float(np.percentile(lats, 99)), 3)},
            "requests": len(lats), "launches": launches, "keys_resident_bytes": stats["resident_bytes"],
            "capacity": capacity,
            "workload": f"{keys} seated keys, Poisson arrivals of 1-64-query batches at uniform tables "
                        f"({len(tables)} resident) and keys at 70% of the batched rate, host inputs "
                        f"(PCIe-inclusive), continuous batching <= {max_launch} queries per launch"}


def batched_leg(g, torch, dev, stream, ctx, seat, shards, B, M, launches, warmup, rng, tune=True, graph=True,
                single=True, trace=True, two_streams=True):
    """Launches of M batches of B seeded queries merged (L = B M queries per launch) over the
    context's `shards` stored shards, back to back on one stream, inputs resident in HBM."""
    L = B * M
    ws = g.Workspace(ctx, L)
    total = warmup + launches
    ids, seeds, bs = make_batches(rng, total * M, B, shards, g.Q)
    d_ids = torch.from_numpy(ids.astype(np.int32)).to(dev).reshape(total, L)
    d_seeds = torch.from_numpy(seeds).to(dev).reshape(total, L, 32)
    d_b = torch.from_numpy(bs.view(np.int64)).to(dev).reshape(total, L, N)
    d_a_out = torch.empty((L, N), dtype=torch.int64, device=dev)
    d_b_out = torch.empty((L, GAMMA), dtype=torch.int64, device=dev)
    torch.cuda.synchronize()

    def issue(i):
        ctx.respond_batch_device(seat, ws, d_ids[i].data_ptr(), d_seeds[i].data_ptr(), d_b[i].data_ptr(), L,
                                 d_a_out.data_ptr(), d_b_out.data_ptr(), stream=stream, switch_modulus=Q36)

    out = {"queries_per_launch": L, "merge": M, "batch": B, "launches": launches, "shards": shards}
    if tune:
        # launch shape picked by timing the three shapes on synthetic queries at this launch size
        out["autotune"] = g.autotune(ctx, seat, ws, L)
        # and the DB scan's row splits (its tile), at the same launch size
        out["autotune"].update(g.autotune_scan(ctx, seat, ws, L))
        log(f"bench: autotune at {L} queries/launch over {shards} shards: {out['autotune']}")
    for i in range(warmup):
        issue(i)
    torch.cuda.synchronize()
    st_ = ws.status(L)
    assert (st_ == 0).all(), f"respond status {st_}"

    def timed_launches(launch):
        starts = [torch.cuda.Event(enable_timing=True) for _ in range(launches)]
        ends = [torch.cuda.Event(enable_timing=True) for _ in range(launches)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(launches):
            starts[k].record()
            launch(k)
            ends[k].record()
        torch.cuda.synchronize()
        wall_ = time.perf_counter() - t0
        st = ws.status(L)
        assert (st == 0).all(), f"respond status {st}"
        return wall_, np.array([s_.elapsed_time(e) for s_, e in zip(starts, ends)])

    # the stream launches of the timed region record events around their scan and chain kernels
    # (no waits); the kernel durations of the roofline are these
    if trace:
        ws.trace(launches)
    wall, lat = timed_launches(lambda k: issue(warmup + k))
    if trace:
        group_ms, scan_ms, chain_ms = ws.kernel_times()
        assert len(chain_ms) == launches, "a timed launch was not traced"
        out["group_ms"], out["scan_ms"] = float(np.mean(group_ms)), float(np.mean(scan_ms))
        out["chain_ms"] = float(np.mean(chain_ms))
        # distinct shards per timed launch: the records one grouped scan reads
        out["distinct_shards"] = float(np.mean([len(np.unique(ids.reshape(total, L)[warmup + k]))
                                                for k in range(launches)]))
    out["stream"] = {"value": round(L * launches / wall, 2), "unit": "queries/s", "wall_s": wall,
                     "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                                    "p99": round(float(np.percentile(lat, 99)), 3)}}
    out["pipeline"], out["value"], out["wall_s"], out["latency"] = "stream", L * launches / wall, wall, lat
    eager_last = d_a_out.clone()
    if two_streams:
        # the same launches alternating between two streams, each with its own workspace and
        # outputs: a launch's grouping, scan and first chains run while the previous launch's
        # last chains drain (one launch in flight per stream, as the serving engine keeps them)
        ws2 = [ws, g.Workspace(ctx, L)]
        outs2 = [(d_a_out, d_b_out), (torch.empty_like(d_a_out), torch.empty_like(d_b_out))]
        st2 = [torch.cuda.Stream(), torch.cuda.Stream()]

        def issue2(i):
            k = i % 2
            ctx.respond_batch_device(seat, ws2[k], d_ids[i].data_ptr(), d_seeds[i].data_ptr(), d_b[i].data_ptr(), L,
                                     outs2[k][0].data_ptr(), outs2[k][1].data_ptr(), stream=st2[k].cuda_stream,
                                     switch_modulus=Q36)

        for i in range(warmup):
            issue2(i)
        torch.cuda.synchronize()
        s2 = [torch.cuda.Event(enable_timing=True) for _ in range(launches)]
        e2 = [torch.cuda.Event(enable_timing=True) for _ in range(launches)]
        t0 = time.perf_counter()
        for k in range(launches):
            s2[k].record(st2[(warmup + k) % 2])  # the stream issue2(warmup + k) launches on
            issue2(warmup + k)
            e2[k].record(st2[(warmup + k) % 2])
        torch.cuda.synchronize()
        wall2 = time.perf_counter() - t0
        for w in ws2:
            st = w.status(L)
            assert (st == 0).all(), f"respond status {st}"
        last = warmup + launches - 1  # the launch whose outputs eager_last holds
        assert torch.equal(outs2[last % 2][0], eager_last), "two-stream launch differs from the stream launch"
        lat2 = np.array([a_.elapsed_time(b_) for a_, b_ in zip(s2, e2)])
        out["two_streams"] = {"value": round(L * launches / wall2, 2), "unit": "queries/s", "wall_s": wall2,
                              "latency_ms": {"p50": round(float(np.percentile(lat2, 50)), 3),
                                             "p99": round(float(np.percentile(lat2, 99)), 3),
                                             "unit": "one launch on its stream, overlapping the other stream's"}}
        if wall2 < wall:
            out["pipeline"], out["value"], out["wall_s"], out["latency"] = "two streams", L * launches / wall2, wall2, lat2
        ws2[1].close()
    if graph:
        graphs = [g.RespondGraph(ctx, seat, ws, d_ids[warmup + k].data_ptr(), d_seeds[warmup + k].data_ptr(),
                                 d_b[warmup + k].data_ptr(), L, d_a_out.data_ptr(), d_b_out.data_ptr(),
                                 switch_modulus=Q36) for k in range(launches)]
        for gr in graphs:
            gr.launch(stream)
        torch.cuda.synchronize()
        d_a_out.zero_()
        gwall, glat = timed_launches(lambda k: graphs[k].launch(stream))
        assert torch.equal(d_a_out, eager_last), "graph replay differs from the stream launch"
        for gr in graphs:
            gr.close()
        out["graph"] = {"value": round(L * launches / gwall, 2), "unit": "queries/s",
                        "latency_ms": {"p50": round(float(np.percentile(glat, 50)), 3),
                                       "p99": round(float(np.percentile(glat, 99)), 3)}}
        # the fastest launch path is the pipeline a server would run; all are reported
        if gwall < out["wall_s"]:
            out["pipeline"], out["value"], out["wall_s"], out["latency"] = "hipGraph", L * launches / gwall, gwall, glat
    if single and M > 1:
        # the same queries answered one B-query batch per launch (no merging across batches)
        nb = min(launches * M, 32)
        t_starts = [torch.cuda.Event(enable_timing=True) for _ in range(nb)]
        t_ends = [torch.cuda.Event(enable_timing=True) for _ in range(nb)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(nb):
            sl, off = warmup + k // M, (k % M) * B
            t_starts[k].record()
            ctx.respond_batch_device(seat, ws, d_ids[sl, off:].data_ptr(), d_seeds[sl, off:].data_ptr(),
                                     d_b[sl, off:].data_ptr(), B, d_a_out.data_ptr(), d_b_out.data_ptr(),
                                     stream=stream, switch_modulus=Q36)
            t_ends[k].record()
        torch.cuda.synchronize()
        sw = time.perf_counter() - t0
        slat = np.array([s_.elapsed_time(e) for s_, e in zip(t_starts, t_ends)])
        out["one_batch_per_launch"] = {"value": round(B * nb / sw, 2), "unit": "queries/s", "batches": nb,
                                       "latency_ms": {"p50": round(float(np.percentile(slat, 50)), 3),
                                                      "p99": round(float(np.percentile(slat, 99)), 3)}}
    ws.close()
    return out


def plain_leg(g, torch, dev, stream, ctx, seat, shards, count, rng):
    """One seeded query per launch, waited for before the next (the plain drop-in)."""
    wsp = g.Workspace(ctx, 1)
    pids, pseeds, pbs = make_batches(rng, count, 1, shards, g.Q)
    pd_ids = torch.from_numpy(pids.astype(np.int32)).to(dev)
    pd_seeds = torch.from_numpy(pseeds).to(dev)
    pd_b = torch.from_numpy(pbs.view(np.int64)).to(dev)
    pa = torch.empty((1, N), dtype=torch.int64, device=dev)
    pb = torch.empty((1, GAMMA), dtype=torch.int64, device=dev)

    def one(i):
        ctx.respond_batch_device(seat, wsp, pd_ids[i].data_ptr(), pd_seeds[i].data_ptr(), pd_b[i].data_ptr(), 1,
                                 pa.data_ptr(), pb.data_ptr(), stream=stream, switch_modulus=Q36)
        torch.cuda.synchronize()

    for i in range(min(2, count)):
        one(i)
    lat = []
    t0 = time.perf_counter()
    for i in range(count):
        t1 = time.perf_counter()
        one(i)
        lat.append((time.perf_counter() - t1) * 1e3)
    qps = count / (time.perf_counter() - t0)
    wsp.close()
    return {"value": round(qps, 2), "unit": "queries/s", "queries": count,
            "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                           "p99": round(float(np.percentile(lat, 99)), 3)}}


# 60-bit modular products of one exact-gadget query at gamma = 256 as the reference's algorithm
# counts them (algorithmic work, not this kernel's own count): NTT(a) (11,264), step 1 (256 x
# 2048), 255 recursion steps of one inverse transform (13,312 with its n^-1 scaling), two forward
# transforms (2 x 11,264), the last digit's derivation (3 x 2048) and the key-switch rows
# (3 x 2048), the online gather (255 x 3 x 2048), and the two final inverse transforms
MM_PER_QUERY = 11264 + 256 * 2048 + 255 * (13312 + 2 * 11264 + 6 * 2048) + 255 * 3 * 2048 + 2 * 13312
assert MM_PER_QUERY == 14_401_536


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16, help="timed launches of the headline (512 queries each)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--merge", type=int, default=8, help="64-query batches merged into one launch")
    ap.add_argument("--shards", type=int, default=512, help="headline shards (512 = 2^20 records of 512 B)")
    ap.add_argument("--c2-shards", type=int, default=128, help="shards of the 2^18-record batched config")
    ap.add_argument("--plain-shards", type=int, default=32, help="shards of the 2^16-record plain config")
    ap.add_argument("--plain-queries", type=int, default=16)
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="time stream launches only")
    ap.add_argument("--no-two-streams", action="store_true",
                    help="time launches on one stream only (no launches alternating between two streams)")
    ap.add_argument("--no-autotune", action="store_true", help="keep --queries-per-cu instead of autotuning")
    ap.add_argument("--no-extra", action="store_true", help="headline only: no 2^18, 2^16 or plain legs")
    ap.add_argument("--served-seconds", type=float, default=4.0)
    ap.add_argument("--queries-per-cu", type=int, default=2, help="launch shaping of the packing kernel (1, 2 or 3)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--cpu-rounds", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sm-tables", type=int, default=2048, help="special-modulus step-1 tables (2048 = 2^22 cells)")
    ap.add_argument("--sm-keys", type=int, default=256)
    ap.add_argument("--sm-batch", type=int, default=2560,
                    help="special-modulus queries per launch (2560: two rounds of 5 per CU on 256 CUs)")
    ap.add_argument("--sm-steps", type=int, default=6)
    ap.add_argument("--sm-only", action="store_true", help="run only the special-modulus workload")
    ap.add_argument("--no-sm", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first copy a 256 MiB device buffer once: a kernel of known HBM traffic that "
                         "scales FETCH_SIZE / WRITE_SIZE counters read from the same profiled run")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the respond path shards by shard id with no exchange, "
              "and this build runs on a single MI355X; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    # HIP maps a process's streams onto GPU_MAX_HW_QUEUES hardware queues (4 unless set), and
    # streams that share a queue run their kernels one after another. The two-stream pipeline
    # and the serving engines (a compute stream per launch slot, two copy streams, the
    # context's stream) need queues of their own: 8, set before HIP starts.
    os.environ.setdefault("GPU_MAX_HW_QUEUES", "8")
    import torch
    import inspire_gpu as g

    assert torch.cuda.is_available(), "no GPU"
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    rng = np.random.default_rng(12345)
    w_seed = rng.integers(0, 256, 32, dtype=np.uint8).tobytes()

    if args.calibrate_copy:
        src = torch.empty(256 << 20, dtype=torch.uint8, device=dev).fill_(1)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    if args.sm_only:
        log("bench: special-modulus workload only")
        sm = special_modulus(g, torch, dev, stream, args)
        if not args.no_cpu:
            sm["cpu_baseline"] = sm_cpu_baseline(args.cpu_seconds)
        print(json.dumps({"special_modulus": sm}), flush=True)
        return

    y = rng.integers(0, g.Q, (3, N), dtype=np.uint64)
    B, M = args.batch, args.merge
    L = B * M

    # --- headline: 2^20 x 512 B, 512 queries per launch
    log(f"bench: building {args.shards} shards of {N} x {ENTRY} B with stores")
    ctx = g.Context(g.params_d2048(), w_seed, GAMMA)
    ctx.set_queries_per_cu(args.queries_per_cu)
    t0 = time.perf_counter()
    build_db(g, ctx, args.shards, seed=1)
    build_s = time.perf_counter() - t0
    seat = ctx.seat(y)
    log("bench: headline launches")
    head = batched_leg(g, torch, dev, stream, ctx, seat, args.shards, B, M, args.steps, args.warmup, rng,
                       tune=not args.no_autotune, graph=not args.no_graph, single=not args.no_extra,
                       two_streams=not args.no_two_streams)
    tuned = head.get("autotune")
    plain = None
    if not args.no_extra:
        log("bench: plain one-query-per-launch path over the same database")
        plain = plain_leg(g, torch, dev, stream, ctx, seat, args.shards, args.plain_queries, rng)
    seat.close()
    ctx.close()

    c2 = c1 = None
    if not args.no_extra:
        log(f"bench: 2^18-record batched config ({args.c2_shards} shards)")
        ctx2 = g.Context(g.params_d2048(), w_seed, GAMMA)
        build_db(g, ctx2, args.c2_shards, seed=3)
        ctx2.set_queries_per_cu(tuned["queries_per_cu"] if tuned else args.queries_per_cu)
        seat2 = ctx2.seat(y)
        c2 = batched_leg(g, torch, dev, stream, ctx2, seat2, args.c2_shards, B, M, min(args.steps, 12), 1, rng,
                         tune=False, graph=False, single=True, trace=False)
        seat2.close()
        ctx2.close()
        log(f"bench: 2^16-record plain config ({args.plain_shards} shards)")
        ctx1 = g.Context(g.params_d2048(), w_seed, GAMMA)
        build_db(g, ctx1, args.plain_shards, seed=2)
        seat1 = ctx1.seat(y)
        c1 = plain_leg(g, torch, dev, stream, ctx1, seat1, args.plain_shards, args.plain_queries, rng)
        seat1.close()
        ctx1.close()

    qps = head["value"]
    launches = args.steps
    scan_avg, chain_avg, group_avg = head["scan_ms"], head["chain_ms"], head["group_ms"]
    # scan: algorithmic bytes per launch = each distinct shard's records (d x entry) and tag key
    # (d x 8) once, and every query's b (d x 8): the shard-grouped scan reads a shard once for all
    # of its queries
    distinct = head["distinct_shards"]
    scan_bytes = distinct * (N * ENTRY + N * 8) + L * N * 8
    scan_gbs = scan_bytes / (scan_avg * 1e-3) / 1e9
    # the same per query (each query's shard counted on its own), as a query-by-query scan reads
    per_query_bytes = L * (N * ENTRY + N * 8 + N * 8)
    chain_gmms = L * MM_PER_QUERY / (chain_avg * 1e-3) / 1e9
    chain_gbs = L * GAMMA * N * 16 / (chain_avg * 1e-3) / 1e9

    # HBM bytes per launch measured by rocprofv3 counters for this workload, committed beside the
    # profile; scaled to this run's launch size
    traffic_chain = traffic_scan = chain_insts = None
    tpath = os.path.join(ROOT, "profiles", "respond_batched512_d2048_g256_2p20_traffic.json")
    if os.path.exists(tpath):
        tj = json.load(open(tpath))
        traffic_chain = round(tj["k_chain_512_queries_bytes"] * L / 512)
        traffic_scan = round(tj["k_scan_512_queries_bytes"] * L / 512)
        if "k_chain_512_queries_valu_insts" in tj:
            chain_insts = tj["k_chain_512_queries_valu_insts"] * L / 512

    sm = None
    if not args.no_sm:
        log("bench: special-modulus workload")
        sm = special_modulus(g, torch, dev, stream, args)

    serve = None
    if not args.no_served:
        log("bench: served config")
        serve = served(g, args.served_seconds, rate_qps=0.7 * qps,
                       capacity_queries=int(max(args.served_seconds, 1.0) * qps),
                       queries_per_cu=tuned["queries_per_cu"] if tuned else args.queries_per_cu)

    cpu = None
    if not args.no_cpu:
        log("bench: CPU baseline")
        cpu = cpu_baseline(args.cpu_seconds, rounds=args.cpu_rounds)
        if sm is not None:
            sm["cpu_baseline"] = sm_cpu_baseline(min(args.cpu_seconds, 10.0))

    lat = head["latency"]
    result = {
        "metric": "respond queries/s + DB GB/s (byte-exact) vs plain GPU; p50/p99 latency",
        "value": round(qps, 2),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": launches,
        "warmup": args.warmup,
        "ms_per_step": round(head["wall_s"] * 1000 / launches, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 (exact mod 2^60-2^14+1, on 32-bit lanes)",
        "data": "synthetic: uniform random records, seeds, query b and key rows from fixed seeds",
        "config": {
            "workload": f"batched TwoPacking respond, d2048 exact gadget, 2^20 records x 512 B ({args.shards} shards, "
                        f"gamma 256), stores precomputed, {B} seeded queries per batch under one seat, uniform "
                        f"indices, 36-bit modulus switch; a step is one launch of {M} merged batches ({L} queries), "
                        f"launches back to back ({head['pipeline']}: "
                        + ("alternating between two streams with a workspace each, one launch in flight per stream"
                           if head["pipeline"] == "two streams" else "one stream") + ")",
            "records": args.shards * N,
            "record_bytes": ENTRY,
            "queries_per_step": L,
            "merge": M,
            "global_batch": B,
            "pipeline": head["pipeline"],
            "queries_per_cu": tuned["queries_per_cu"] if tuned else args.queries_per_cu,
            "autotune": tuned,
        },
        "db_gbs": round(qps * N * ENTRY / 1e9, 3),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                       "p99": round(float(np.percentile(lat, 99)), 3),
                       "unit": f"one launch ({L} queries), events on its stream"
                               + (", overlapping the other stream's launch" if head["pipeline"] == "two streams"
                                  else "")},
        "one_batch_per_launch": head.get("one_batch_per_launch"),
        "graph_replays": head.get("graph"),
        "stream_launches": head["stream"],
        "two_stream_launches": head.get("two_streams"),
        "plain_gpu": None if plain is None else dict(plain, workload="one seeded query per launch over the same "
                                                                      "2^20-record database, waited for in turn"),
        "vs_plain_gpu": None if plain is None else round(qps / plain["value"], 3),
        "batched_2p18": None if c2 is None else {
            "workload": f"2^18 records x 512 B ({args.c2_shards} shards), {L} queries per launch and one "
                        f"{B}-query batch per launch, stream launches",
            "value": c2["stream"]["value"], "unit": "queries/s", "latency_ms": c2["stream"]["latency_ms"],
            "one_batch_per_launch": c2.get("one_batch_per_launch")},
        "plain_2p16": None if c1 is None else dict(c1, workload=f"one seeded query per launch, 2^16 records x 512 B "
                                                                 f"({args.plain_shards} shards)"),
        "roofline": {"bound": "valu", "kernel": "k_chain",
                     "achieved": round(chain_gmms * 4 / 1e3, 3), "peak": VALU_MAD_PEAK_T,
                     "unit": "T mad32/s", "frac": round(chain_gmms * 4 / 1e3 / VALU_MAD_PEAK_T, 4),
                     "traffic": traffic_chain,
                     "launch_ms": round(chain_avg, 4), "mm_per_query": MM_PER_QUERY,
                     "timing": "HIP events on the launch stream around this kernel in each launch of the timed region",
                     "modmul60_per_s": round(chain_gmms * 1e9, 0),
                     "definition": "4 v_mad_u64_u32 (32x32->64) per 60-bit modular product, algorithmic; peak = "
                                   "256 CU x 64 lanes x 2.4 GHz (v_mad_u64_u32 issues at half the simple-op rate)",
                     "valu_issue": valu_issue(chain_insts, chain_avg,
                                              "profiles/respond_batched512_d2048_g256_2p20.md")},
        "roofline_scan": {"bound": "hbm", "kernel": "k_scan_grouped", "achieved": round(scan_gbs, 2),
                          "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(scan_gbs / HBM_PEAK_GBS, 4),
                          "traffic": traffic_scan, "launch_ms": round(scan_avg, 4),
                          "algorithmic_bytes": round(scan_bytes), "distinct_shards_per_launch": distinct,
                          "grouping_kernel_ms": round(group_avg, 4),
                          "per_query_bytes": per_query_bytes,
                          "per_query_gbs": round(per_query_bytes / ((scan_avg + group_avg) * 1e-3) / 1e9, 2),
                          "timing": "HIP events on the launch stream around this kernel in each launch of the timed "
                                    "region, over the 2^20-record database (512 MiB of records); per_query_gbs "
                                    "divides the bytes a query-by-query scan would read by the grouping and scan "
                                    "kernels' time"},
        "store_stream_gbs": round(chain_gbs, 2),
        "build_s": round(build_s, 2),
        "served": serve,
        "special_modulus": sm,
        "cpu_baseline": cpu,
    }
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
