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
              its HIP-event-timed duration on the launch stream, against 8 TB/s HBM;
  cpu_baseline the reference's own CPU matmul (oracle/_ref) on a bounded sample of the
              same workload, one host thread.

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


def cpu_baseline(arrays, E, budget_s=12.0):
    """The reference CPU matmul (oracle/_ref, or the oracle's faithful port if the reference
    library is absent) on a bounded sample: f32 K=16 traversals, UP and DOWN alternately."""
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
    done = 0
    t0 = time.perf_counter()
    while True:
        run(v, oracle.UP)
        run(w, oracle.DOWN)
        done += 2
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    return {
        "value": E * K * done / dt,
        "unit": "edge*RHS/s",
        "cores": 1,
        "kind": kind,
        "sample": f"{done} single-thread f32 K=16 traversals (UP and DOWN alternating) of the same graph, "
                  f"{dt:.1f} s",
    }


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
    cg = capi.Graph(arrays)
    info = cg.info()
    V, E, N, M = info["num_nodes"], info["num_edges"], info["num_samples"], info["num_mutations"]
    log(f"graph: V={V} E={E} levels up/down={info['up_levels']}/{info['down_levels']} "
        f"max fan-in up/down={info['max_fanin_up']}/{info['max_fanin_down']} built in {time.time() - t0:.1f}s")
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
    if not args.no_autotune:
        for cdt, d, _, _, _ in work:
            best, ms = cg.autotune(d, cdt, Kpad)
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
    # Timed region: level-kernel HIP events on the launch stream for the f32 traversals.
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        step(timing_f32=True)
    torch.cuda.synchronize()
    t_total = time.perf_counter() - t_start
    k_ms, k_launches, k_bytes = cg.kernel_timing_read(reset=True)
    ms_per_step = t_total / args.steps * 1e3
    edge_rhs_per_step = 4 * E * Ktot
    value = edge_rhs_per_step / (t_total / args.steps)
    step_alg_bytes = sum(algorithmic_bytes(V, E, Ktot, 4, cin, cout) for _, _, cin, cout, _ in work)
    log(f"batched step: {ms_per_step:.3f} ms, {value / 1e9:.1f} G edge*RHS/s, "
        f"{step_alg_bytes / (t_total / args.steps) / 1e9:.0f} GB/s algorithmic")

    # Plain path: one request per traversal, f32, each K of the ladder.
    plain = {}
    plain_ms_total = 0.0
    for cdt, d, cin, cout, reqs in work:
        if cdt != capi.F32 or args.no_plain:
            continue
        name = "UP" if d == capi.UP else "DOWN"
        for K, x, y in reqs:
            if not args.no_autotune:
                cg.autotune(d, cdt, K)
            reps = 5
            cg.matmul_ptr(cdt, x.data_ptr(), cin, K, d, y.data_ptr(), y.numel(), stream=sp)
            torch.cuda.synchronize()
            t = time.perf_counter()
            for _ in range(reps):
                cg.matmul_ptr(cdt, x.data_ptr(), cin, K, d, y.data_ptr(), y.numel(), stream=sp)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t) / reps * 1e3
            plain_ms_total += ms
            plain[f"{name}_K{K}"] = {
                "ms": round(ms, 4),
                "edge_rhs_per_s": E * K / (ms / 1e3),
                "alg_GBps": algorithmic_bytes(V, E, K, 4, cin, cout) / (ms / 1e3) / 1e9,
            }
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

    achieved = k_bytes / (k_ms / 1e3) / 1e9 if k_ms > 0 else 0.0
    roofline = {
        "bound": "hbm",
        "kernel": "pull_level_kernel<float,16,*,false> (level gather, f32, 511 stacked RHS)",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic_from_profile(),
        "launches": k_launches,
        "avg_launch_us": round(k_ms / max(1, k_launches) * 1e3, 2),
    }
    cpu = None
    if not args.no_cpu_baseline:
        log("timing the CPU baseline ...")
        cpu = cpu_baseline(arrays, E)
    result = {
        "metric": "GRG matmul edge*RHS/s (batched K-ladder, UP+DOWN, f32+i32)",
        "value": value,
        "unit": "edge*RHS/s",
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
        "step_alg_GBps": step_alg_bytes / (t_total / args.steps) / 1e9,
        "plain": plain,
        "batched_vs_plain_f32": round(plain_ms_total / batched_f32_ms, 3) if plain_ms_total else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(result), flush=True)
    cg.close()


if __name__ == "__main__":
    main()
