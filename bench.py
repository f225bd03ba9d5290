"""Benchmark: batched GRG matrix multiplication on one MI355X.

Workload (the batched right-hand-side configuration): a synthetic GRG of 200,000
haplotypes x 1,000,000 biallelic sites built from 50 seeded Kingman coalescent trees
(one per genome segment), resident on the device.  One step = four traversals, each
carrying the whole K ladder {1, 2, 4, ..., 256} stacked as 511 right-hand sides sharing
one pass over the graph: UP and DOWN in float32 (N(0,1) inputs) and UP and DOWN in int32
(Bernoulli(0.5) masks).  Inputs and outputs are device-resident before timing starts.

value = edge*RHS per second over the step (graph edges x stacked rows / step time).

Also reported:
  plain       one request per traversal for each K of the ladder (f32, UP and DOWN), the
              per-K edge*RHS/s and algorithmic GB/s, and the batched/plain speedup;
  roofline    the float32 level (pull-gather) kernel: algorithmic bytes per launch over
              its HIP-event-timed duration on the launch stream, against 8 TB/s HBM; beside
              it the rocprofv3-measured HBM bytes per launch over the same duration and the
              device's measured streaming read / copy rates (grgl_hip_stream_ceiling);
  cpu_baseline the reference's own CPU matmul (oracle/_ref) on a bounded sample of the
              same workload: all host cores (one independent call per thread) as value,
              one thread beside it;
  fp32_error  per result (graph, K, direction): FP32 error against the FP64 oracle, the ratio
              to the a priori rounding bound, and the reference's own FP32 error beside it
              (checked after the timed region);
  serving     burst / Poisson / small-K streams, served vs one request per traversal, with a
              checked sample of served results compared with the same requests alone.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# The headline metric this benchmark reports (edge*RHS/s of the batched traversal, with
# the HBM roofline fraction in "roofline" and served p50/p99 vs the plain path in "serving").
METRIC = "GRG matmul edge\u00b7RHS/s + HBM GB/s frac; served p50/p99 vs plain, 1 GPU"
K_LADDER = [1, 2, 4, 8, 16, 32, 64, 128, 256]
SEED = 20261203


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes(V, E, K, es, cin, cout):
    """4(E+V) index bytes + s*K*(E + V + 2*C_in + 2*C_out) value bytes per traversal."""
    return 4 * (E + V) + es * K * (E + V + 2 * cin + 2 * cout)


def traffic_from_profile():
    """HBM bytes per level-kernel launch measured with rocprofv3 --pmc, if recorded."""
    path = os.path.join(ROOT, "profiles", "pull_level_traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(arrays, E, budget_s=8.0):
    """The reference CPU matmul (oracle/_ref, or the oracle's faithful port if the reference
    library is absent) on a bounded sample: f32 K=16 traversals, UP and DOWN alternately.

    Two legs, as the reference is used: (1) one thread, back-to-back calls (the reference has
    no intra-call threading); (2) all host cores available to this process (at most 16), one
    independent call per thread on the shared immutable graph.  `value` is the all-cores
    throughput; the single-thread figure is reported beside it."""
    import concurrent.futures as cf

    import numpy as np

    import oracle

    og = oracle.OracleGraph(arrays)
    kind = "reference" if oracle.ref_available() else "port"
    rng = np.random.default_rng(SEED)
    K = 16
    v = rng.standard_normal((K, og.num_samples)).astype(np.float32)
    w = rng.standard_normal((K, og.num_mutations)).astype(np.float32)
    if kind == "reference":
        og.ref_handle()  # build the reference CSRGRG (varint encoding) outside the timing

        def run(x, d):
            return og.ref_matmul(x, d)
    else:
        def run(x, d):
            return og.faithful_matmul(x, d)

    def worker(deadline):
        n = 0
        while True:
            run(v, oracle.UP)
            run(w, oracle.DOWN)
            n += 2
            if time.perf_counter() >= deadline:
                return n

    t0 = time.perf_counter()
    done1 = worker(t0 + budget_s)
    dt1 = time.perf_counter() - t0
    single = E * K * done1 / dt1
    try:
        threads = min(16, len(os.sched_getaffinity(0)))
    except AttributeError:
        threads = min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(threads) as ex:
        counts = list(ex.map(worker, [t0 + budget_s] * threads))
    dtn = time.perf_counter() - t0
    done = sum(counts)
    # Single-thread latency per call at K = 1, 16 and 256 (UP and DOWN), one call each after
    # a warm-up for K = 256 (about a second per call), a few for the smaller K.
    latency = {}
    for KL, reps in ((1, 5), (16, 2), (256, 1)):
        vl = rng.standard_normal((KL, og.num_samples)).astype(np.float32)
        wl = rng.standard_normal((KL, og.num_mutations)).astype(np.float32)
        latency[f"K{KL}"] = {"UP_ms": round(_cpu_ms(og, kind, vl, oracle.UP, reps), 2),
                             "DOWN_ms": round(_cpu_ms(og, kind, wl, oracle.DOWN, reps), 2)}
    return {
        "value": E * K * done / dtn,
        "unit": "edge\u00b7RHS/s",
        "cores": threads,
        "kind": kind,
        "single_thread_value": single,
        "single_thread_latency": latency,
        "host_cpus": os.cpu_count(),
        "cpu_model": _cpu_model(),
        "sample": f"f32 K=16 traversals (UP and DOWN alternating) of the benchmark graph: {done} calls on "
                  f"{threads} threads in {dtn:.1f} s (independent calls, shared graph); single thread: "
                  f"{done1} calls in {dt1:.1f} s",
    }


SERVE_GRAPHS = [(10_000, 50_000, 10), (100_000, 500_000, 20), (200_000, 1_000_000, 50)]
SERVE_K = [1, 8, 32, 128]
SERVE_P = [0.4, 0.3, 0.2, 0.1]


def serve_benchmark(num_requests=2000, load=0.6, num_streams=2):
    """Resident graphs, mixed requests: continuous batching vs one request per traversal.

    Requests pick a graph uniformly, K from {1,8,32,128} (p .4/.3/.2/.1), a direction
    uniformly and f32 (70%) or i32 (30%).  Inputs and outputs are device buffers (one
    input/output pair per graph, direction and dtype, sized for K=128; every request
    reads rows [0, K) and writes rows [0, K) of them, so concurrent requests of one key
    write identical values).  Scenarios:
      burst   all requests queued at once: throughput (requests/s, edge*RHS/s);
      poisson arrivals at `load` x the plain path's burst throughput: p50/p99 latency
              (arrival -> result written), measured by the server for the served path and
              by the caller around each synchronous call for the plain path.
    """
    import numpy as np
    import torch

    import grgl_amd
    from grgl_amd import capi

    sp = torch.cuda.current_stream().cuda_stream
    rng = np.random.default_rng(SEED + 5)
    graphs = []
    for n, m, t in SERVE_GRAPHS:
        g = grgl_amd.grg_from_coalescent(n, m, t, seed=SEED + n)
        cg = capi.Graph(g.arrays())
        graphs.append(cg)
    bufs = {}
    for gi, cg in enumerate(graphs):
        for d in (capi.UP, capi.DOWN):
            cin, cout = (cg.num_samples, cg.num_mutations) if d == capi.UP else (cg.num_mutations, cg.num_samples)
            for cdt, tdt in ((capi.F32, torch.float32), (capi.I32, torch.int32)):
                x = torch.randn((128, cin), device="cuda").to(tdt)
                y = torch.zeros((128, cout), device="cuda", dtype=tdt)
                bufs[(gi, d, cdt)] = (x, y, cin, cout)
    edges = [cg.info()["num_edges"] for cg in graphs]
    reqs = []
    for _ in range(num_requests):
        gi = int(rng.integers(0, len(graphs)))
        K = int(rng.choice(SERVE_K, p=SERVE_P))
        d = capi.UP if rng.random() < 0.5 else capi.DOWN
        cdt = capi.F32 if rng.random() < 0.7 else capi.I32
        reqs.append((gi, K, d, cdt))
    edge_rhs = sum(edges[gi] * K for gi, K, _, _ in reqs)

    def plain_one(r):
        gi, K, d, cdt = r
        x, y, cin, cout = bufs[(gi, d, cdt)]
        graphs[gi].matmul_ptr(cdt, x.data_ptr(), cin, K, d, y.data_ptr(), K * cout, stream=sp)

    # Warm every (graph, direction, dtype, K) shape once on both paths.
    for gi in range(len(graphs)):
        for d in (capi.UP, capi.DOWN):
            for cdt in (capi.F32, capi.I32):
                for K in SERVE_K:
                    plain_one((gi, K, d, cdt))
    torch.cuda.synchronize()
    srv = capi.Server(num_streams=num_streams, max_batch_rows=512, hbm_budget_bytes=48 << 30)
    gids = [srv.add_graph(cg) for cg in graphs]

    # A checked sample of the stream (every 25th request) writes to outputs of its own, so its
    # served results can be compared with the same requests run alone after the timing.
    private = {}
    for i in range(0, num_requests, 25):
        gi, K, d, cdt = reqs[i]
        x, y, cin, cout = bufs[(gi, d, cdt)]
        private[i] = torch.zeros((K, cout), device="cuda", dtype=y.dtype)

    def served_submit(r, i=None):
        gi, K, d, cdt = r
        x, y, cin, cout = bufs[(gi, d, cdt)]
        if i in private:
            y = private[i]
        return srv.submit_device(gids[gi], cdt, d, K, x.data_ptr(), y.data_ptr())

    for r in reqs[:64]:
        srv.wait(served_submit(r))
    batches0 = srv.stats(reset=True)["batches"]

    # Burst: plain.
    torch.cuda.synchronize()
    t = time.perf_counter()
    for r in reqs:
        plain_one(r)
    torch.cuda.synchronize()
    plain_burst = time.perf_counter() - t
    # Burst: served.
    t = time.perf_counter()
    ids = [served_submit(r, i) for i, r in enumerate(reqs)]
    for rid in ids:
        srv.wait(rid)
    served_burst = time.perf_counter() - t
    st_burst = srv.stats(reset=True)
    batches1 = st_burst["batches"]
    # The checked sample against the same requests alone (bit for bit).
    checked_equal = True
    for i, y_served in private.items():
        gi, K, d, cdt = reqs[i]
        x, _, cin, cout = bufs[(gi, d, cdt)]
        alone = torch.zeros_like(y_served)
        graphs[gi].matmul_ptr(cdt, x.data_ptr(), cin, K, d, alone.data_ptr(), alone.numel(), stream=sp)
        torch.cuda.synchronize()
        checked_equal = checked_equal and bool(torch.equal(alone.view(torch.int32), y_served.view(torch.int32)))
    plain_rate = num_requests / plain_burst
    rate = load * plain_rate
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=num_requests))

    def spin_until(t0, at):
        while True:
            now = time.perf_counter() - t0
            if now >= at:
                return
            if at - now > 2e-4:
                time.sleep((at - now) / 2)

    # Poisson: plain (FIFO, one at a time).
    lat = []
    t0 = time.perf_counter()
    for r, at in zip(reqs, arrivals):
        spin_until(t0, at)
        plain_one(r)
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - t0 - at) * 1e3)
    plain_poisson_wall = time.perf_counter() - t0
    # Poisson: served.
    t0 = time.perf_counter()
    ids = []
    for r, at in zip(reqs, arrivals):
        spin_until(t0, at)
        ids.append(served_submit(r))
    for rid in ids:
        srv.wait(rid)
    served_poisson_wall = time.perf_counter() - t0
    st = srv.stats(reset=True)
    # Burst of latency-bound requests only (K in {1, 8}): where batching pays most.
    small = []
    for _ in range(num_requests):
        gi = int(rng.integers(0, len(graphs)))
        K = int(rng.choice([1, 8]))
        d = capi.UP if rng.random() < 0.5 else capi.DOWN
        cdt = capi.F32 if rng.random() < 0.7 else capi.I32
        small.append((gi, K, d, cdt))
    small_edge_rhs = sum(edges[gi] * K for gi, K, _, _ in small)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for r in small:
        plain_one(r)
    torch.cuda.synchronize()
    plain_small = time.perf_counter() - t
    t = time.perf_counter()
    ids = [served_submit(r) for r in small]
    for rid in ids:
        srv.wait(rid)
    served_small = time.perf_counter() - t
    st_small = srv.stats(reset=True)
    srv.close()
    for cg in graphs:
        cg.close()
    lat = np.array(lat)
    return {
        "workload": "3 resident graphs (10k x 50k / 100k x 500k / 200k x 1M), K in {1,8,32,128} p .4/.3/.2/.1, "
                    "direction uniform, 70% f32 / 30% i32, device-resident request buffers",
        "requests": num_requests,
        "burst": {
            "plain_req_per_s": round(plain_rate, 1),
            "served_req_per_s": round(num_requests / served_burst, 1),
            "plain_edge_rhs_per_s": edge_rhs / plain_burst,
            "served_edge_rhs_per_s": edge_rhs / served_burst,
            "speedup": round(plain_burst / served_burst, 3),
            "served_batches": batches1 - batches0,
            "served_p50_ms": round(st_burst["latency_p50_ms"], 3),
            "served_p99_ms": round(st_burst["latency_p99_ms"], 3),
            "checked_requests": len(private),
            "checked_equal_alone": checked_equal,
        },
        "poisson": {
            "offered_req_per_s": round(rate, 1),
            "plain_p50_ms": round(float(np.percentile(lat, 50)), 3),
            "plain_p99_ms": round(float(np.percentile(lat, 99)), 3),
            "served_p50_ms": round(st["latency_p50_ms"], 3),
            "served_p99_ms": round(st["latency_p99_ms"], 3),
            "plain_wall_s": round(plain_poisson_wall, 3),
            "served_wall_s": round(served_poisson_wall, 3),
            "served_batches": st["batches"] - batches1,
        },
        "burst_small_k": {
            "workload": "same graphs and mix of direction / dtype, K in {1, 8} p .5/.5",
            "plain_req_per_s": round(num_requests / plain_small, 1),
            "served_req_per_s": round(num_requests / served_small, 1),
            "plain_edge_rhs_per_s": small_edge_rhs / plain_small,
            "served_edge_rhs_per_s": small_edge_rhs / served_small,
            "speedup": round(plain_small / served_small, 3),
            "served_batches": st_small["batches"] - st["batches"],
        },
    }


def options_benchmark(cg, N, M, V, E, P=2, K=16):
    """The remaining matmul options on the benchmark graph, f32, K = 16, one request per call:
    by_individual (UP reads K x N/P, DOWN writes K x N/P), emit_all_nodes (K x V out) and
    order_preserving (every list summed in the reference's order, heavy rows included)."""
    import torch

    from grgl_amd import capi

    sp = torch.cuda.current_stream().cuda_stream
    cases = [("by_individual_UP", capi.UP, N // P, M, dict(by_individual=True)),
             ("by_individual_DOWN", capi.DOWN, M, N // P, dict(by_individual=True)),
             ("emit_all_nodes_UP", capi.UP, N, V, dict(emit_all=True)),
             ("emit_all_nodes_DOWN", capi.DOWN, M, V, dict(emit_all=True)),
             ("order_preserving_UP", capi.UP, N, M, dict(flags=capi.FLAG_ORDER_PRESERVING)),
             ("order_preserving_DOWN", capi.DOWN, M, N, dict(flags=capi.FLAG_ORDER_PRESERVING))]
    out = {}
    for name, d, cin, cout, kw in cases:
        x = torch.randn((K, cin), device="cuda")
        y = torch.zeros((K, cout), device="cuda")
        for _ in range(2):
            cg.matmul_ptr(capi.F32, x.data_ptr(), cin, K, d, y.data_ptr(), y.numel(), stream=sp, **kw)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(5):
            cg.matmul_ptr(capi.F32, x.data_ptr(), cin, K, d, y.data_ptr(), y.numel(), stream=sp, **kw)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t) / 5 * 1e3
        out[name] = {"ms": round(ms, 4), "edge_rhs_per_s": E * K / (ms / 1e3)}
    return out


def host_path_benchmark(g, E):
    """pygrgl-style calls with host numpy input and output (grgl_amd.matmul), f32: the time per
    call including the staging copies over PCIe, beside the device-resident traversal."""
    import numpy as np

    import grgl_amd

    rng = np.random.default_rng(SEED)
    out = {}
    for d, name, cin, cout in ((grgl_amd.UP, "UP", g.num_samples, g.num_mutations),
                               (grgl_amd.DOWN, "DOWN", g.num_mutations, g.num_samples)):
        for K in (1, 16, 256):
            x = rng.standard_normal((K, cin)).astype(np.float32)
            grgl_amd.matmul(g, x, d)
            reps = 5 if K < 256 else 3
            t = time.perf_counter()
            for _ in range(reps):
                grgl_amd.matmul(g, x, d)
            ms = (time.perf_counter() - t) / reps * 1e3
            moved = 4 * K * (cin + cout)
            out[f"{name}_K{K}"] = {"ms": round(ms, 3), "edge_rhs_per_s": E * K / (ms / 1e3),
                                   "host_bytes": moved, "host_GBps": round(moved / (ms / 1e3) / 1e9, 2)}
    return out


def grg_io_benchmark(g):
    """Write and read the benchmark graph as a .grg file (format 5.3) in a temporary directory."""
    import tempfile

    import grgl_amd

    with tempfile.TemporaryDirectory() as tmp:
        fn = os.path.join(tmp, "bench.grg")
        t = time.perf_counter()
        grgl_amd.save_grg(g, fn)
        w = time.perf_counter() - t
        size = os.path.getsize(fn)
        t = time.perf_counter()
        h = grgl_amd.load_immutable_grg(fn)
        r = time.perf_counter() - t
        same = (h.num_nodes, h.num_edges, h.num_mutations) == (g.num_nodes, g.num_edges, g.num_mutations)
    return {"file_bytes": size, "write_s": round(w, 3), "read_s": round(r, 3),
            "read_MBps": round(size / r / 1e6, 1), "round_trip_equal": bool(same)}


def levelized_file_benchmark(arrays, cg, create_s):
    """The benchmark graph's levelized-graph file (grgl_hip_graph_store / grgl_hip_graph_load,
    the reference's storeGPUGRGToDisk / loadGPUGRGFromDisk): write time, size, and load time
    beside grgl_hip_graph_create's (levelize on the host + upload); the loaded graph's int32
    UP K = 4 result must equal the created graph's."""
    import tempfile

    import numpy as np
    import torch

    from grgl_amd import capi

    out = {"graph_create_s": round(create_s, 3)}
    with tempfile.TemporaryDirectory() as tmp:
        fn = os.path.join(tmp, "bench.lvlz")
        t = time.perf_counter()
        cg.store(fn)
        out["store_s"] = round(time.perf_counter() - t, 3)
        out["file_bytes"] = os.path.getsize(fn)
        t = time.perf_counter()
        capi.levelize_to_file(arrays, fn + ".host")
        out["levelize_to_file_s"] = round(time.perf_counter() - t, 3)
        t = time.perf_counter()
        loaded = capi.Graph.load(fn)
        out["load_s"] = round(time.perf_counter() - t, 3)
    sp = torch.cuda.current_stream().cuda_stream
    x = (torch.rand((4, cg.num_samples), device="cuda") < 0.5).to(torch.int32)
    ys = []
    for g in (cg, loaded):
        y = torch.zeros((4, g.num_mutations), dtype=torch.int32, device="cuda")
        g.matmul_ptr(capi.I32, x.data_ptr(), g.num_samples, 4, capi.UP, y.data_ptr(), y.numel(), stream=sp)
        torch.cuda.synchronize()
        ys.append(y.cpu().numpy())
    out["loaded_equals_created"] = bool(np.array_equal(ys[0], ys[1]))
    loaded.close()
    return out


def construction_benchmark():
    """Graph construction from the carrier lists of the reference's test VCF (400 haplotypes x
    10,893 sites): map_mutations (smallest carrier sets first, one mutation per batch, one host
    core) and the haplotype-clustering constructor (fast_grg_from_samples, default options)."""
    import numpy as np

    import grgl_amd

    fx = np.load(os.path.join(ROOT, "tests", "golden", "test200.npz"))
    off, car = fx["carrier_off"], fx["carriers"]
    sets = [car[off[m]:off[m + 1]].astype(np.uint32).tolist() for m in range(len(off) - 1)]
    muts = [grgl_amd.Mutation(float(p), "A", "G") for p in fx["positions"]]
    g = grgl_amd.MutableGRG(int(fx["num_samples"]), int(fx["ploidy"]))
    t = time.perf_counter()
    grgl_amd.map_mutations(g, muts, sets, mutation_batch_size=1, size_order=True)
    dt = time.perf_counter() - t
    entries = int(sum(len(x) for x in sets))
    t = time.perf_counter()
    fb = grgl_amd.fast_grg_from_samples(int(fx["num_samples"]), int(fx["ploidy"]), carrier_offsets=off, carriers=car,
                                        positions=fx["positions"])
    fb_s = time.perf_counter() - t
    return {"workload": "test-200-samples VCF carriers: 400 haplotypes x 10,893 sites",
            "seconds": round(dt, 3), "mutations_per_s": round(len(sets) / dt, 1),
            "nodes": g.num_nodes, "edges": g.num_edges, "carrier_entries": entries,
            "edges_per_carrier": round(g.num_edges / entries, 4),
            "fast_grg_from_samples": {"seconds": round(fb_s, 3), "nodes": fb.num_nodes, "edges": fb.num_edges,
                                      "edges_per_carrier": round(fb.num_edges / entries, 4)}}


def constructed_graph_benchmark(haplotypes=100_000, sites=500_000, trees=20, autotune=True,
                                graphs=("tree_direct", "constructed")):
    """The traversal on the graph shape the haplotype-clustering constructor builds
    (fast_grg_from_samples), beside the tree-direct graph of the same genotype matrix:
    `haplotypes` x `sites` from `trees` seeded coalescent segment trees, carriers taken from the
    tree-direct graph, the constructed graph saved and reloaded (simplified on write, as its
    .grg file holds it).  Per graph: shape (nodes, edges, levels, fan-in histogram), plain f32
    ms at K = 1, 16, 256 with the default launch shapes and autotuned, and the K ladder
    (511 stacked RHS) in one traversal with the level kernels' live HIP-event time.  Checked
    after timing: the int32 products of the two graphs are bit-identical (one genotype matrix),
    and the constructed graph's f32 K = 1 and 16 results are within 1e-6 normwise of the FP64
    oracle.  Default size: the 100,000 x 500,000 (20 trees) workload."""
    import tempfile

    import numpy as np
    import torch

    import grgl_amd
    from grgl_amd import capi

    import oracle
    from oracle import fp_bounds

    t = time.perf_counter()
    src = grgl_amd.grg_from_coalescent(haplotypes, sites, trees, seed=SEED)
    off, car = src.mutation_carriers()
    log(f"constructed graph: {len(car)} carrier entries in {time.perf_counter() - t:.1f} s; building ...")
    t = time.perf_counter()
    built, st = grgl_amd.fast_grg_from_samples(haplotypes, 2, carrier_offsets=off, carriers=car,
                                               positions=src.arrays()["positions"], return_stats=True)
    build_s = time.perf_counter() - t
    log(f"constructed graph: built in {build_s:.1f} s ({st.trees} trees)")
    with tempfile.TemporaryDirectory() as tmp:
        fn = os.path.join(tmp, "constructed.grg")
        grgl_amd.save_grg(built, fn)
        con = grgl_amd.load_immutable_grg(fn)
    del built
    out = {"workload": f"{haplotypes:,} haplotypes x {sites:,} sites ({trees} coalescent segment trees); the "
                       "constructed graph built from the carrier lists",
           "carrier_entries": int(len(car)), "build_s": round(build_s, 2),
           "build": {"trees": st.trees, "windows": st.windows, "unique_hap_segments": st.unique_hap_segments,
                     "queries": st.queries, "comparisons": st.comparisons,
                     "truncated_queries": st.truncated_queries, "combined_nodes": st.combined_nodes}}
    del off, car
    sp = torch.cuda.current_stream().cuda_stream
    gen = torch.Generator(device="cuda")
    gen.manual_seed(SEED + 3)
    checks = {}
    f32_checks = []
    for name, g in (("tree_direct", src), ("constructed", con)):
        if name not in graphs:
            continue
        a = g.arrays()
        if name == "constructed":
            con_arrays = a
        cg = capi.Graph(a)
        info = cg.info()
        V, E, N, M = info["num_nodes"], info["num_edges"], info["num_samples"], info["num_mutations"]
        shape = g.graph_stats()
        rec = {"nodes": V, "edges": E, "levels_up": info["up_levels"], "levels_down": info["down_levels"],
               "max_fanin_up": info["max_fanin_up"], "max_fanin_down": info["max_fanin_down"],
               "fanin_histogram_log2": shape["fan_in_histogram"]}
        for d, dn in ((capi.UP, "UP"), (capi.DOWN, "DOWN")):
            cin, cout = (N, M) if d == capi.UP else (M, N)
            for K in (1, 16, 256):
                x = torch.randn((K, cin), device="cuda", generator=gen)
                y = torch.zeros((K, cout), device="cuda")

                def timed():
                    cg.matmul_ptr(capi.F32, x.data_ptr(), cin, K, d, y.data_ptr(), y.numel(), stream=sp)
                    torch.cuda.synchronize()
                    reps = 10
                    t0 = time.perf_counter()
                    for _ in range(reps):
                        cg.matmul_ptr(capi.F32, x.data_ptr(), cin, K, d, y.data_ptr(), y.numel(), stream=sp)
                    torch.cuda.synchronize()
                    return (time.perf_counter() - t0) / reps * 1e3

                ms_default = timed()
                r = {"ms_default": round(ms_default, 4)}
                if autotune:
                    best, _ = cg.autotune(d, capi.F32, K)
                    ms_tuned = timed()
                    r.update({"ms_tuned": round(ms_tuned, 4), "tuning": best})
                ms = min(ms_default, r.get("ms_tuned", ms_default))
                r["edge_rhs_per_s"] = E * K / (ms / 1e3)
                r["alg_GBps"] = round(algorithmic_bytes(V, E, K, 4, cin, cout) / (ms / 1e3) / 1e9, 1)
                rec[f"plain_f32_{dn}_K{K}"] = r
                if name == "constructed" and K in (1, 16):
                    f32_checks.append((dn, K, x.cpu().numpy(), y.cpu().numpy()))
            # The K ladder stacked in one traversal, level kernels timed with HIP events.
            Ktot = sum(K_LADDER)
            x = torch.randn((Ktot, cin), device="cuda", generator=gen)
            y = torch.zeros((Ktot, cout), device="cuda")
            reqs, k0 = [], 0
            for K in K_LADDER:
                reqs.append((K, x[k0:k0 + K].data_ptr(), y[k0:k0 + K].data_ptr()))
                k0 += K
            if autotune:
                cg.autotune(d, capi.F32, Ktot)
            cg.matmul_batch_ptr(capi.F32, d, reqs, stream=sp)
            torch.cuda.synchronize()
            cg.kernel_timing_read(reset=True)
            reps = 5
            t0 = time.perf_counter()
            for _ in range(reps):
                cg.kernel_timing(True)
                cg.matmul_batch_ptr(capi.F32, d, reqs, stream=sp)
                cg.kernel_timing(False)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) / reps * 1e3
            k_ms, k_n, k_bytes = cg.kernel_timing_read(reset=True)
            achieved = k_bytes / (k_ms / 1e3) / 1e9 if k_ms > 0 else 0.0
            rec[f"batched_f32_{dn}_511rhs"] = {
                "ms": round(ms, 4), "edge_rhs_per_s": E * Ktot / (ms / 1e3),
                "alg_GBps": round(algorithmic_bytes(V, E, Ktot, 4, cin, cout) / (ms / 1e3) / 1e9, 1),
                "level_kernels_GBps": round(achieved, 1), "level_kernels_frac": round(achieved / HBM_PEAK_GBS, 4),
                "level_launches": k_n // reps, "level_ms_per_traversal": round(k_ms / reps, 4),
                "level_alg_bytes_per_traversal": int(k_bytes / reps)}
            # int32 products for the cross-graph check
            xi = (torch.rand((16, cin), device="cuda", generator=torch.Generator(device="cuda").manual_seed(
                SEED + (1 if d == capi.UP else 2))) < 0.5).to(torch.int32)
            yi = torch.zeros((16, cout), device="cuda", dtype=torch.int32)
            cg.matmul_ptr(capi.I32, xi.data_ptr(), cin, 16, d, yi.data_ptr(), yi.numel(), stream=sp)
            torch.cuda.synchronize()
            checks.setdefault(dn, []).append(yi.cpu().numpy())
        cg.close()
        out[name] = rec
        log(f"{name}: {json.dumps(rec)}")
    if len(graphs) == 2:
        out["int32_products_identical"] = bool(all(np.array_equal(v[0], v[1]) for v in checks.values()))
    if "constructed" not in graphs:
        return out
    og = oracle.OracleGraph(con_arrays)
    errs = []
    for dn, K, x, y in f32_checks:
        rep = fp_bounds.fp32_report(og, x, oracle.UP if dn == "UP" else oracle.DOWN, y, with_reference=False)
        errs.append({"direction": dn, "K": K, "normwise": float("%.4g" % rep["normwise"]),
                     "bound_ratio": round(rep["bound_ratio"], 4)})
    out["constructed_fp32_error"] = errs
    return out


def fp32_errors(arrays3, f32_results):
    """FP32 error of the traversal results, stated per result (a checker leg, run after the
    timed region): against the FP64 oracle (normwise per row, max over rows), against the a
    priori rounding bound of a correct FP32 evaluation (bound_ratio <= 1), and the same
    normwise error of the reference's own FP32 result (oracle/_ref) beside it.
      1,000 x 10,000 (1 tree) and 100,000 x 500,000 (20 trees): one request per traversal,
        K = 1 and 16, UP and DOWN, N(0,1) inputs;
      200,000 x 1,000,000 (50 trees): the K = 1 and K = 16 requests of the timed 511-RHS
        batched traversal (`f32_results`: direction -> [(K, x, y)])."""
    import numpy as np
    import torch

    import grgl_amd
    import oracle
    from grgl_amd import capi
    from oracle import fp_bounds

    sp = torch.cuda.current_stream().cuda_stream
    out = []

    def record(name, path, K, d, x, y, og):
        rep = fp_bounds.fp32_report(og, x, oracle.UP if d == capi.UP else oracle.DOWN, y)
        rep.update({"graph": name, "path": path, "K": K, "direction": "UP" if d == capi.UP else "DOWN"})
        for k in ("normwise", "ref_fp32_normwise"):
            rep[k] = float("%.4g" % rep[k])
        rep["bound_ratio"] = round(rep["bound_ratio"], 4)
        out.append(rep)

    rng = np.random.default_rng(SEED + 7)
    for n, m, t in ((1000, 10_000, 1), (100_000, 500_000, 20)):
        a = grgl_amd.grg_from_coalescent(n, m, t, seed=SEED).arrays()
        cg = capi.Graph(a)
        og = oracle.OracleGraph(a)
        for d in (capi.UP, capi.DOWN):
            cin, cout = (n, m) if d == capi.UP else (m, n)
            for K in (1, 16):
                x = rng.standard_normal((K, cin)).astype(np.float32)
                xd = torch.from_numpy(x).cuda()
                yd = torch.zeros((K, cout), device="cuda")
                cg.matmul_ptr(capi.F32, xd.data_ptr(), cin, K, d, yd.data_ptr(), yd.numel(), stream=sp)
                torch.cuda.synchronize()
                record(f"{n} x {m}", "plain", K, d, x, yd.cpu().numpy(), og)
        cg.close()
    og3 = oracle.OracleGraph(arrays3)
    for d, reqs in f32_results.items():
        for K, x, y in reqs:
            if K in (1, 16):
                record("200000 x 1000000", "batched 511 RHS", K, d, x.cpu().numpy(), y.cpu().numpy(), og3)
    return {"bar": "normwise <= 1e-6 against FP64 on these graphs; bound_ratio <= 1",
            "max_normwise": max(r["normwise"] for r in out), "max_bound_ratio": max(r["bound_ratio"] for r in out),
            "results": out}


def _cpu_ms(og, kind, x, direction, reps):
    """Single-thread milliseconds per call of the CPU matmul (the reference's own, oracle/_ref,
    or the oracle's faithful port)."""
    run = og.ref_matmul if kind == "reference" else og.faithful_matmul
    if reps > 1:
        run(x, direction)  # warm-up (page-in); a single long call is timed as it comes
    t = time.perf_counter()
    for _ in range(reps):
        run(x, direction)
    return (time.perf_counter() - t) / reps * 1e3


def plain_at_scale(with_cpu=True):
    """One request per traversal on 100,000 haplotypes x 500,000 sites (20 trees):
    K = 1 and 16, FP64 and FP32, UP and DOWN (ms per traversal, edge*RHS/s, algorithmic GB/s),
    with the CPU matmul's single-thread milliseconds per call on the same input beside it."""
    import numpy as np
    import torch

    import grgl_amd
    from grgl_amd import capi

    g = grgl_amd.grg_from_coalescent(100_000, 500_000, 20, seed=SEED)
    arrays = g.arrays()
    cg = capi.Graph(arrays)
    og = kind = None
    if with_cpu:
        import oracle

        og = oracle.OracleGraph(arrays)
        kind = "reference" if oracle.ref_available() else "port"
        if kind == "reference":
            og.ref_handle()
    info = cg.info()
    V, E, N, M = info["num_nodes"], info["num_edges"], info["num_samples"], info["num_mutations"]
    sp = torch.cuda.current_stream().cuda_stream
    out = {"workload": "100,000 haplotypes x 500,000 sites, 20 trees", "nodes": V, "edges": E}
    for tdt, cdt, es, name in ((torch.float64, capi.F64, 8, "f64"), (torch.float32, capi.F32, 4, "f32")):
        for d, dn in ((capi.UP, "UP"), (capi.DOWN, "DOWN")):
            cin, cout = (N, M) if d == capi.UP else (M, N)
            for K in (1, 16):
                x = torch.randn((K, cin), device="cuda", dtype=tdt)
                y = torch.zeros((K, cout), device="cuda", dtype=tdt)
                cg.autotune(d, cdt, K)
                cg.matmul_ptr(cdt, x.data_ptr(), cin, K, d, y.data_ptr(), y.numel(), stream=sp)
                torch.cuda.synchronize()
                reps = 20
                t = time.perf_counter()
                for _ in range(reps):
                    cg.matmul_ptr(cdt, x.data_ptr(), cin, K, d, y.data_ptr(), y.numel(), stream=sp)
                torch.cuda.synchronize()
                ms = (time.perf_counter() - t) / reps * 1e3
                rec = {"ms": round(ms, 4), "edge_rhs_per_s": E * K / (ms / 1e3),
                       "alg_GBps": round(algorithmic_bytes(V, E, K, es, cin, cout) / (ms / 1e3) / 1e9, 1)}
                if og is not None:
                    xh = np.ascontiguousarray(x.cpu().numpy())
                    od = oracle.UP if d == capi.UP else oracle.DOWN
                    rec["cpu_1thread_ms"] = round(_cpu_ms(og, kind, xh, od, 2), 2)
                out[f"{name}_{dn}_K{K}"] = rec
    if og is not None:
        out["cpu_kind"] = kind
    cg.close()
    return out


def pca_benchmark():
    """Standardized PCA: 50,000 diploid individuals (100,000 haplotypes) x 300,000 sites,
    15 segment trees; randomized SVD, block 20 (10 + 10 oversampling), 5 power iterations;
    FP32 traversals, FP64 statistics and QR on the GPU.  Errors against the FP64 CPU oracle
    running the same algorithm with the same random block."""
    import numpy as np
    import torch

    import grgl_amd
    from grgl_amd import capi
    from grgl_amd.pca import DeviceStdOperator, eigen_errors, randomized_pca

    g = grgl_amd.grg_from_coalescent(100_000, 300_000, 15, seed=SEED + 4)
    arrays = g.arrays()
    cg = capi.Graph(arrays)
    op = DeviceStdOperator(cg)
    randomized_pca(op, k=10)  # warm-up
    torch.cuda.synchronize()
    t = time.perf_counter()
    e, U = randomized_pca(op, k=10)
    torch.cuda.synchronize()
    gpu_s = time.perf_counter() - t
    out = {"workload": "50,000 diploids (100,000 haplotypes) x 300,000 sites, 15 trees; k=10, block 20, "
                       "5 power iterations", "gpu_ms": round(gpu_s * 1e3, 2), "traversals": 12,
           "top10_eigenvalues": [float(x) for x in e]}
    try:
        from oracle.pca_oracle import OracleStdOperator
        from oracle.pca_oracle import randomized_pca as oracle_pca

        t = time.perf_counter()
        omega = op.randn(op.shape[1], 20, 20261203).cpu().numpy()  # the block the device run drew
        re, rU = oracle_pca(OracleStdOperator(arrays), k=10, omega=omega)
        out["cpu_oracle_fp64_s"] = round(time.perf_counter() - t, 2)
        rel, ang = eigen_errors(e, U, re, rU)
        out["max_eigenvalue_rel_err"] = float(rel.max())
        out["max_eigenvector_angle_rad"] = float(ang.max())
    except ImportError:
        pass
    cg.close()
    return out


def latency_trace(args):
    """One request per traversal, FP32, K = 1, 4, 16, 64, DOWN then UP per K, library defaults:
    3 untimed calls, then 20 timed calls per shape; one JSON line of milliseconds per call."""
    import torch

    import grgl_amd
    from grgl_amd import capi

    g = grgl_amd.grg_from_coalescent(args.haplotypes, args.sites, args.trees, seed=SEED)
    cg = capi.Graph(g.arrays())
    sp = torch.cuda.current_stream().cuda_stream
    N, M = g.num_samples, g.num_mutations
    out = {"workload": f"{args.haplotypes} haplotypes x {args.sites} sites, {args.trees} trees; FP32, one request "
                       "per traversal", "ms": {}}
    for K in (1, 4, 16, 64):
        for d, dn, cin, cout in ((capi.DOWN, "DOWN", M, N), (capi.UP, "UP", N, M)):
            x = torch.randn((K, cin), device="cuda")
            y = torch.zeros((K, cout), device="cuda")
            for _ in range(3):
                cg.matmul_ptr(capi.F32, x.data_ptr(), cin, K, d, y.data_ptr(), y.numel(), stream=sp)
            torch.cuda.synchronize()
            t = time.perf_counter()
            for _ in range(20):
                cg.matmul_ptr(capi.F32, x.data_ptr(), cin, K, d, y.data_ptr(), y.numel(), stream=sp)
            torch.cuda.synchronize()
            out["ms"][f"{dn}_K{K}"] = round((time.perf_counter() - t) / 20 * 1e3, 4)
            log(f"{dn} K={K}: {out['ms'][f'{dn}_K{K}']} ms")
            time.sleep(0.3)
    cg.close()
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--haplotypes", type=int, default=200_000)
    ap.add_argument("--sites", type=int, default=1_000_000)
    ap.add_argument("--trees", type=int, default=50)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the one-request-per-traversal sweep")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving (mixed request stream) scenario")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--no-pca", action="store_true", help="skip the standardized PCA scenario")
    ap.add_argument("--no-fp32-error", action="store_true",
                    help="skip the FP32 error statement (oracle comparisons after the timed region)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the option, .grg file and construction measurements")
    ap.add_argument("--latency-trace", action="store_true",
                    help="only time one-request FP32 traversals of the benchmark graph at K = 1, 4, 16 (UP and "
                         "DOWN, library default tuning), 20 calls per shape with 0.3 s between shapes, so a "
                         "kernel trace splits by shape")
    ap.add_argument("--constructed-only", action="store_true",
                    help="only measure the constructed-graph shape beside the tree-direct graph")
    ap.add_argument("--constructed-graphs", default="tree_direct,constructed",
                    help="which graphs the constructed-graph comparison measures (for profiling one of them)")
    ap.add_argument("--constructed-size", default="100000,500000,20",
                    help="haplotypes,sites,trees of the constructed-graph comparison (default: 100,000 x 500,000, "
                         "20 trees)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch

        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 4, "copy_bytes_written": dst.numel() * 4}))
        return
    if args.latency_trace:
        latency_trace(args)
        return
    if args.constructed_only:
        n, m, t = (int(v) for v in args.constructed_size.split(","))
        print(json.dumps(constructed_graph_benchmark(n, m, t, autotune=not args.no_autotune,
                                                     graphs=tuple(args.constructed_graphs.split(",")))), flush=True)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: the batched GRG traversal runs on a single MI355X here; "
              "multi-GPU sharding (by RHS columns or genome segment) is not part of this build.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import grgl_amd
    from grgl_amd import capi

    torch.cuda.set_device(0)
    t0 = time.time()
    g = grgl_amd.grg_from_coalescent(args.haplotypes, args.sites, args.trees, seed=SEED)
    arrays = g.arrays()
    t_create = time.time()
    cg = capi.Graph(arrays)  # levelization of both directions + upload (grgl_hip_graph_create)
    graph_create_s = time.time() - t_create
    info = cg.info()
    V, E, N, M = info["num_nodes"], info["num_edges"], info["num_samples"], info["num_mutations"]
    log(f"graph: V={V} E={E} levels up/down={info['up_levels']}/{info['down_levels']} "
        f"max fan-in up/down={info['max_fanin_up']}/{info['max_fanin_down']} built in {time.time() - t0:.1f}s "
        f"(device graph {graph_create_s:.2f}s)")
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    gen = torch.Generator(device="cuda")
    gen.manual_seed(SEED)
    Ktot = sum(K_LADDER)

    def make(dtype, rows, cols):
        if dtype == torch.float32:
            return torch.randn((rows, cols), device="cuda", generator=gen)
        return (torch.rand((rows, cols), device="cuda", generator=gen) < 0.5).to(torch.int32)

    # Stacked requests: one tensor per K, per (dtype, direction).
    work = []
    for tdt, cdt in ((torch.float32, capi.F32), (torch.int32, capi.I32)):
        for d in (capi.UP, capi.DOWN):
            cin, cout = (N, M) if d == capi.UP else (M, N)
            reqs = []
            for K in K_LADDER:
                x = make(tdt, K, cin)
                y = torch.zeros((K, cout), device="cuda", dtype=tdt)
                reqs.append((K, x, y))
            work.append((cdt, d, cin, cout, reqs))
    torch.cuda.synchronize()

    # The library stacks the ladder largest-first on 16-byte boundaries: 511 rows exactly.
    Kpad = Ktot
    tuning_chosen = {}
    if not args.no_autotune:
        for cdt, d, _, _, _ in work:
            best, ms = cg.autotune(d, cdt, Kpad)
            tuning_chosen[f"{'f32' if cdt == capi.F32 else 'i32'}_{'UP' if d == capi.UP else 'DOWN'}_K{Kpad}"] = dict(
                best, tuned_ms=round(ms, 4))
            log(f"autotune dtype={cdt} dir={d} rows={Kpad}: {best} levels {ms:.3f} ms")

    def step(timing_f32=False):
        for cdt, d, cin, cout, reqs in work:
            cg.kernel_timing(timing_f32 and cdt == capi.F32)
            cg.matmul_batch_ptr(cdt, d, [(K, x.data_ptr(), y.data_ptr()) for K, x, y in reqs], stream=sp)
        cg.kernel_timing(False)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    cg.kernel_timing_read(reset=True)
    # Timed region: level-kernel HIP events on the launch stream for the f32 traversals of one
    # step in four (each event pair costs the stream a few microseconds; one step in four keeps
    # that below 1% of the step while every kernel duration is still measured live here).
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        step(timing_f32=(i % 4 == 0))
    torch.cuda.synchronize()
    t_total = time.perf_counter() - t_start
    k_ms, k_launches, k_bytes = cg.kernel_timing_read(reset=True)
    ms_per_step = t_total / args.steps * 1e3
    edge_rhs_per_step = 4 * E * Ktot
    value = edge_rhs_per_step / (t_total / args.steps)
    step_alg_bytes = sum(algorithmic_bytes(V, E, Ktot, 4, cin, cout) for _, _, cin, cout, _ in work)
    log(f"batched step: {ms_per_step:.3f} ms, {value / 1e9:.1f} G edge*RHS/s, "
        f"{step_alg_bytes / (t_total / args.steps) / 1e9:.0f} GB/s algorithmic")

    # Plain path: one request per traversal, each K of the ladder, f32 (and i32 beside it).
    plain = {}
    plain_i32 = {}
    plain_ms_total = 0.0
    for cdt, d, cin, cout, reqs in work:
        if args.no_plain:
            continue
        name = "UP" if d == capi.UP else "DOWN"
        for K, x, y in reqs:
            if not args.no_autotune:
                best, ms = cg.autotune(d, cdt, K)
                tuning_chosen[f"{'f32' if cdt == capi.F32 else 'i32'}_{name}_K{K}"] = dict(best, tuned_ms=round(ms, 4))
            reps = 5
            cg.matmul_ptr(cdt, x.data_ptr(), cin, K, d, y.data_ptr(), y.numel(), stream=sp)
            torch.cuda.synchronize()
            t = time.perf_counter()
            for _ in range(reps):
                cg.matmul_ptr(cdt, x.data_ptr(), cin, K, d, y.data_ptr(), y.numel(), stream=sp)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t) / reps * 1e3
            rec = {
                "ms": round(ms, 4),
                "edge_rhs_per_s": E * K / (ms / 1e3),
                "alg_GBps": algorithmic_bytes(V, E, K, 4, cin, cout) / (ms / 1e3) / 1e9,
            }
            if cdt == capi.F32:
                plain_ms_total += ms
                plain[f"{name}_K{K}"] = rec
            else:
                plain_i32[f"{name}_K{K}"] = rec
    batched_f32_ms = None
    # Batched f32 UP+DOWN alone, to compare with the plain per-K sum on equal work.
    f32_work = [w for w in work if w[0] == capi.F32]
    torch.cuda.synchronize()
    t = time.perf_counter()
    reps = 5
    for _ in range(reps):
        for cdt, d, cin, cout, reqs in f32_work:
            cg.matmul_batch_ptr(cdt, d, [(K, x.data_ptr(), y.data_ptr()) for K, x, y in reqs], stream=sp)
    torch.cuda.synchronize()
    batched_f32_ms = (time.perf_counter() - t) / reps * 1e3
    log(f"plain per-K f32 UP+DOWN total {plain_ms_total:.3f} ms vs batched {batched_f32_ms:.3f} ms")
    fp32_err = None
    if not args.no_fp32_error:
        log("FP32 error against the FP64 oracle ...")
        fp32_err = fp32_errors(arrays, {d: reqs for cdt, d, _, _, reqs in f32_work})
        log(f"fp32 error: max normwise {fp32_err['max_normwise']:.3g}, max bound ratio "
            f"{fp32_err['max_bound_ratio']:.3g}")

    achieved = k_bytes / (k_ms / 1e3) / 1e9 if k_ms > 0 else 0.0
    avg_launch_us = k_ms / max(1, k_launches) * 1e3
    traffic = traffic_from_profile()
    # The practical streaming limit of this GPU (one 16-byte access per thread), measured now.
    ceiling = capi.stream_ceiling()
    log(f"streaming ceiling: read {ceiling['read_GBps']:.0f} GB/s, copy {ceiling['copy_GBps']:.0f} GB/s")
    # The practical random-gather rate (4-byte values, the access of a traversal at K = 1):
    # tables of 2 MB (a K = 1 row buffer of this graph's size, cache resident) and 256 MB.
    gather = {f"table_{tb >> 20}MB_lanes_{lpr}": round(capi.gather_ceiling(table_bytes=tb, lanes_per_row=lpr) / 1e9, 1)
              for tb in (2 << 20, 256 << 20) for lpr in (1, 4)}
    log(f"random-gather ceiling (G gathers/s): {gather}")
    roofline = {
        "bound": "hbm",
        "kernel": "pull_level_kernel<float,16,*,false> (level gather, f32, 511 stacked RHS)",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "launches": k_launches,
        "avg_launch_us": round(avg_launch_us, 2),
        # rocprofv3 HBM bytes per launch over the live average launch time
        "traffic_GBps": round(traffic / (avg_launch_us * 1e3), 1) if traffic and k_launches else None,
        "measured_read_ceiling_GBps": round(ceiling["read_GBps"], 1),
        "measured_copy_ceiling_GBps": round(ceiling["copy_GBps"], 1),
    }
    serving = None
    if not args.no_serving:
        log("serving scenario ...")
        serving = serve_benchmark(args.serve_requests)
        log(f"serving: {json.dumps(serving)}")
    at_scale = None
    if not args.no_plain:
        log("plain path at scale (100k x 500k) ...")
        at_scale = plain_at_scale(with_cpu=not args.no_cpu_baseline)
    pca = None
    if not args.no_pca:
        log("PCA scenario ...")
        pca = pca_benchmark()
        log(f"pca: {json.dumps(pca)}")
    extras = None
    if not args.no_extras:
        log("options, .grg file and construction ...")
        extras = {"options_K16_f32": options_benchmark(cg, N, M, V, E), "host_path_f32": host_path_benchmark(g, E),
                  "grg_io": grg_io_benchmark(g), "construction": construction_benchmark(),
                  "levelized_file": levelized_file_benchmark(arrays, cg, graph_create_s)}
        log("constructed-graph shape ...")
        n, m, t = (int(v) for v in args.constructed_size.split(","))
        extras["constructed_graph"] = constructed_graph_benchmark(n, m, t, autotune=not args.no_autotune)
        log(f"extras: {json.dumps(extras)}")
    cpu = None
    if not args.no_cpu_baseline:
        log("timing the CPU baseline ...")
        cpu = cpu_baseline(arrays, E)
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "edge\u00b7RHS/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+i32",
        "data": "synthetic (seeded Kingman coalescent trees; N(0,1) f32 and Bernoulli(0.5) i32 inputs)",
        "config": {
            "workload": "batched RHS: 200,000 haplotypes x 1,000,000 sites, 50 segment trees; "
                        "K ladder 1..256 stacked (511 RHS) per traversal; UP+DOWN x f32+i32 per step",
            "haplotypes": args.haplotypes,
            "sites": args.sites,
            "trees": args.trees,
            "nodes": V,
            "edges": E,
            "stacked_rhs": Ktot,
            "levels_up": info["up_levels"],
            "levels_down": info["down_levels"],
            "parallelism": "single GPU",
        },
        # host levelization of both directions + upload, once per graph (not in the step)
        "graph_create_s": round(graph_create_s, 3),
        "step_alg_GBps": step_alg_bytes / (t_total / args.steps) / 1e9,
        "plain": plain,
        "plain_i32": plain_i32,
        "gather_ceiling_G_per_s": gather,
        "batched_vs_plain_f32": round(plain_ms_total / batched_f32_ms, 3) if plain_ms_total else None,
        "roofline": roofline,
        # autotuned level-kernel tile per (dtype, direction, stacked rows): team width, block
        # threads, loads in flight, hipGraph replay, block order, chained launch
        "tuning": tuning_chosen,
        "fp32_error": fp32_err,
        "plain_100k_x_500k": at_scale,
        "serving": serving,
        "pca": pca,
        "extras": extras,
        "cpu_baseline": cpu,
    }
    print(json.dumps(result), flush=True)
    cg.close()


if __name__ == "__main__":
    main()
