"""Benchmark: Mini-FEM per-iteration assembly + block-Jacobi preconditioner on one MI355X.

Workload (default): the ela operator on a 100^3-node jittered Kuhn cube (1,000,000 nodes,
5,821,794 tets, 14,761,198 CSR blocks of 3x3), node ids shuffled, per-element Lame
coefficients; geometry seed 3, coefficient seed 103, shuffle seed 203, jitter 0.25h.
A step is one FEM iteration (src/FEM.cc:183-232): assembly of all CSR values plus the
inverted block-Jacobi prec, with all inputs resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--schedule tiled|colored]
                    [--precision fp64|mixed] [--op ela|lap] [--n 100] [--no-cpu]

Prints one JSON line: metric/value (CSR blocks assembled per second), the roofline of
the fused kernel (algorithmic HBM bytes per launch / its mean launch time, against the
8 TB/s HBM3E peak), the host-core baselines (the CPU oracle's REF-order iteration and
its COLORING-order OpenMP iteration on the same mesh, both OpenMP variants, timed here),
the accuracy of the GPU results against that oracle iteration (or, with --no-cpu in
mixed precision, against an FP64 step), and -- for the default ela 100^3 line -- the
other two workloads of the metric: "batched" (64 meshes of 5k-200k nodes in one captured
graph, FP64 and mixed) and "served" (16 resident meshes, a Poisson request stream under an
HBM budget: p50/p99 latency and throughput of plain FP64 one request at a time against
continuous batching in FP64 and mixed precision).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "mini-fem_amd"), os.path.join(ROOT, "oracle")]

HBM_PEAK_GBS = 8000.0
METRIC = "CSR blocks assembled/s & %HBM peak; p50/p99 req latency vs plain FP64 GPU"


def algorithmic_bytes(nE, N, nnz, op):
    """Compulsory HBM bytes of one iteration: connectivity 16/elem, coefficients
    8c/elem, coordinates 24/node, row_ptr, col_idx read once, values and prec written
    once, ela boundary mask 12/node."""
    c, d = (2, 9) if op == 1 else (1, 1)
    b = 16 * nE + 8 * c * nE + 24 * N + 4 * (N + 1) + 4 * nnz + 8 * d * nnz + 8 * d * N
    return b + (12 * N if op == 1 else 0)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def delaunay_arrays(npts, seed=5):
    """An unstructured mesh: the Delaunay tetrahedralization of npts seeded random points in
    the unit cube (nearly flat tets dropped), hull nodes with seeded boundary codes. Returns
    (coord, elemToNode 1-based, boundNodesCode, nnz) like minifem.generate_cube."""
    import numpy as np
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(seed)
    pts = rng.random((npts, 3))
    tri = Delaunay(pts)
    tets = tri.simplices.astype(np.int64)
    a, b, c = (pts[tets[:, k]] - pts[tets[:, 3]] for k in range(3))
    tets = tets[np.abs(np.einsum("ij,ij->i", a, np.cross(b, c))) > 1e-9 / npts]
    i, j = tets[:, [0, 0, 0, 1, 1, 2]].ravel(), tets[:, [1, 2, 3, 2, 3, 3]].ravel()
    nnz = npts + 2 * np.unique(np.minimum(i, j) * npts + np.maximum(i, j)).size
    code = np.zeros(npts, np.int32)
    hull = np.unique(tri.convex_hull)
    code[hull] = rng.choice([10, 52, 53, 54], hull.size)
    return pts.ravel().copy(), (tets + 1).astype(np.int32).ravel(), code, int(nnz)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--schedule", default="tiled", choices=["tiled", "colored"])
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--op", default="ela", choices=["ela", "lap"])
    ap.add_argument("--n", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--delaunay", type=int, default=0,
                    help="instead of the cube: the Delaunay tetrahedralization of this many "
                         "seeded random points in the unit cube (an unstructured mesh)")
    ap.add_argument("--no-layers", action="store_true",
                    help="leave out the batched and served legs of the default line")
    ap.add_argument("--cpu-iters", type=int, default=10,
                    help="iterations of the faster (per-color parallel) CPU baseline; the two "
                         "serial-element-loop baselines run 2 (iteration 0 excluded from all)")
    # schedule of the tiled plan: the library's defaults, tuned for the 100^3 cubes on MI355X
    # (smaller meshes gain ~10 % from --tune, which runs minifem.tune on the mesh first)
    ap.add_argument("--tune", action="store_true",
                    help="pick the tiled schedule with the autotuner on this mesh first")
    ap.add_argument("--tile-lds", type=int, default=155000)
    ap.add_argument("--block", type=int, default=1024)
    ap.add_argument("--max-rows", type=int, default=254)
    ap.add_argument("--id-chunks", type=int, default=1,
                    help="bisect each of this many contiguous node-id ranges on its own")
    ap.add_argument("--batched", action="store_true",
                    help="instead: 64 meshes (5k-200k nodes, lap/ela) in one captured graph")
    ap.add_argument("--batch-forms", default="per_mesh,union",
                    help="batched forms to run: per_mesh, union or both (comma-separated)")
    ap.add_argument("--batch-params", default="",
                    help="LDS,BLOCK,ROWS: fixed batch schedule instead of tuning it")
    ap.add_argument("--serve", action="store_true",
                    help="instead: Poisson request stream on 16 resident meshes (1k-1M nodes)")
    ap.add_argument("--requests", type=int, default=300)
    ap.add_argument("--rate", type=float, default=400.0, help="served arrivals per second")
    ap.add_argument("--write-probe", action="store_true",
                    help="instead of the FEM step, time the step's output writes alone: the "
                         "tile rows in the kernel's order vs the same bytes front to back")
    ap.add_argument("--gather", type=int, default=0,
                    help="instead of the FEM step, time the random gather probe: this many "
                         "8-byte reads at seeded random positions of a 1 GiB array (the byte "
                         "counters' scale for scattered reads)")
    ap.add_argument("--copy", type=int, default=0,
                    help="instead of the FEM step, time the plain HBM copy kernel on this many "
                         "bytes (bandwidth reference and byte-counter calibration)")
    ap.add_argument("--pcg", type=int, default=0,
                    help="instead of the FEM step, time this many iterations of the block-Jacobi "
                         "PCG (mf_pcg) on the operator and preconditioner the step assembles "
                         "for the --op/--n cube")
    ap.add_argument("--coloring", action="store_true",
                    help="instead of the FEM step, time the greedy coloring of the --n cube "
                         "(node ids shuffled) on the GPU against the host scan, same result")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: --gpus {args.gpus} is not supported by this "
              "build (the single-GPU assembly step has no multi-GPU path)", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import minifem
    import minifem.gpu as gpu

    if args.copy:
        gpu.init(0)
        n = (args.copy // 16) * 16
        src = gpu.DeviceArray(n // 8, np.float64)
        dst = gpu.DeviceArray(n // 8, np.float64)
        stream = gpu.Stream()
        lib = minifem.lib()
        for _ in range(args.warmup):
            minifem.check(lib.mf_hbm_copy(dst.ptr, src.ptr, n, stream.ptr), "mf_hbm_copy")
        e0, e1 = gpu.Event(), gpu.Event()
        e0.record(stream)
        for _ in range(args.steps):
            minifem.check(lib.mf_hbm_copy(dst.ptr, src.ptr, n, stream.ptr), "mf_hbm_copy")
        e1.record(stream)
        ms = e0.elapsed_ms(e1) / args.steps
        print(json.dumps({"copy_bytes": n, "ms_per_copy": ms,
                          "gbs_read_plus_write": 2 * n / ms / 1e6}), flush=True)
        return

    if args.gather:
        gpu.init(0)
        span = 1 << 27  # doubles: 1 GiB
        n = int(args.gather)
        rng = np.random.default_rng(7)
        idx = rng.integers(0, span, n, dtype=np.int64)
        src = gpu.DeviceArray(span, np.float64)
        didx = gpu.DeviceArray.from_numpy(idx)
        dst = gpu.DeviceArray(n, np.float64)
        stream = gpu.Stream()
        lib = minifem.lib()
        for _ in range(args.warmup):
            minifem.check(lib.mf_gather_probe(dst.ptr, src.ptr, didx.ptr, n, stream.ptr), "gather")
        e0, e1 = gpu.Event(), gpu.Event()
        e0.record(stream)
        for _ in range(args.steps):
            minifem.check(lib.mf_gather_probe(dst.ptr, src.ptr, didx.ptr, n, stream.ptr), "gather")
        e1.record(stream)
        ms = e0.elapsed_ms(e1) / args.steps
        # distinct 64- and 128-byte lines the gathers touch (what a cold DRAM read moves)
        lines = {w: int(np.unique(idx // (w // 8)).size) for w in (32, 64, 128)}
        print(json.dumps({"gathers": n, "ms": ms, "index_bytes": 8 * n, "dst_bytes": 8 * n,
                          "distinct_lines": lines,
                          "gather_bytes_if_line": {w: lines[w] * w for w in lines}}), flush=True)
        return

    if args.coloring:
        gpu.init(0)
        coord, e2n, _, _ = minifem.generate_cube(args.n, 0.25, 3, 203)
        N = coord.size // 3
        d_e2n = gpu.DeviceArray.from_numpy(np.ascontiguousarray(e2n, np.int32))
        t_dev, t_host = [], []
        for _ in range(max(args.steps, 1)):
            gpu.synchronize()
            t = time.perf_counter()
            part, perm, c2e = gpu.coloring_creation(d_e2n, N)
            t_dev.append(time.perf_counter() - t)
        t = time.perf_counter()
        hp, hperm, hc2e = minifem.coloring_creation(e2n, N)
        t_host.append(time.perf_counter() - t)
        nE = e2n.size // 4
        same = bool(np.array_equal(part.numpy()[:nE], hp) and np.array_equal(perm.numpy()[:nE], hperm)
                    and np.array_equal(c2e, hc2e))
        print(json.dumps({"coloring": f"cube {args.n}^3 node-shuffled, {nE} tets, {len(hc2e) - 1} colors",
                          "device_ms_median": round(float(np.median(t_dev)) * 1e3, 2),
                          "device_ms_first": round(t_dev[0] * 1e3, 2),
                          "host_ms": round(t_host[0] * 1e3, 2), "equal": same}), flush=True)
        return

    if args.pcg:
        return bench_pcg(args)
    if args.batched:
        return bench_batched(args)
    if args.serve:
        return bench_serve(args)

    op = minifem.OP_ELA if args.op == "ela" else minifem.OP_LAP
    prec_mode = minifem.PREC_FP64 if args.precision == "fp64" else minifem.PREC_MIXED
    gpu.init(0)
    t0 = time.time()
    if args.delaunay:
        coord, e2n, code, nnz = delaunay_arrays(args.delaunay)
        shape = f"Delaunay mesh of {args.delaunay} random points"
    else:
        coord, e2n, code, nnz = minifem.generate_cube(args.n, 0.25, 3, 203)
        shape = f"{args.n}^3 jittered Kuhn cube, node-shuffled"
    coef = minifem.generate_coef(op, e2n.size // 4, 103)
    mesh = minifem.Mesh(coord, e2n, code, nnz, op, coef=coef)
    log(f"setup: {mesh.nbNodes} nodes, {mesh.nbElem} tets, {mesh.nnz} blocks, "
        f"{mesh.nbColors} colors ({time.time() - t0:.1f}s)")
    dm = gpu.DeviceMesh(mesh)
    plan = None
    tuning = None
    if args.schedule == "tiled":
        # a mixed step's tiles are sized by its float stage
        params = dict(tile_lds_bytes=args.tile_lds, block_threads=args.block,
                      max_tile_rows=args.max_rows, id_chunks=args.id_chunks,
                      stage_precision=prec_mode)
        if args.tune:
            from minifem.tune import tune
            params, res = tune(mesh, dm, precision=prec_mode, save=False, log=log)
            tuning = [dict(c, us=round(us, 2), tiles=st["tiles"]) for c, us, st in res]
        t1 = time.time()
        plan = gpu.Plan(mesh, **params)
        log(f"plan: {params} {plan.stats()} ({time.time() - t1:.1f}s)")
    stream = gpu.Stream()

    out_bytes = 8 * mesh.nnz * (9 if op == minifem.OP_ELA else 1)

    def probe_us(mode, steps):
        """Mean time of the plan's output stream alone (mf_plan_write_probe)."""
        for _ in range(args.warmup):
            plan.write_probe(dm.values, mode, stream)
        e0, e1 = gpu.Event(), gpu.Event()
        e0.record(stream)
        for _ in range(steps):
            plan.write_probe(dm.values, mode, stream)
        e1.record(stream)
        return e0.elapsed_ms(e1) / steps * 1e3

    if args.write_probe:
        res = {}
        for mode, name in ((0, "tile_rows"), (1, "front_to_back"), (2, "tile_rows_whole_lines"),
                           (3, "tile_rows_partial_lines")):
            us = probe_us(mode, args.steps)
            res[name] = {"us": round(us, 1)}
            if mode < 2:  # modes 2 and 3 write a part of the bytes each
                res[name]["TB_per_s"] = round(out_bytes / us / 1e6, 3)
        print(json.dumps({"write_probe": res, "bytes": out_bytes,
                          "workload": f"{args.op} {shape}, tiled plan "
                                      f"{plan.stats()['tiles']} tiles"}), flush=True)
        return

    # the floor the values' own write order sets, measured before the timed steps (the
    # probe overwrites the values; the warm-up steps assemble them again)
    floor = None
    if plan is not None:
        floor = {"tile_rows_us": round(probe_us(0, 20), 1),
                 "front_to_back_us": round(probe_us(1, 20), 1), "bytes": out_bytes}

    def step(precision=prec_mode):
        if plan is not None:
            plan.assemble(dm.coord, dm.coef, dm.values, dm.prec, precision=precision,
                          stream=stream)
        else:
            dm.iterate_colored(precision=precision, stream=stream)

    for _ in range(args.warmup):
        step()
    stream.sync()
    K = args.steps
    ev = [gpu.Event() for _ in range(K + 1)]
    gpu.synchronize()
    t_wall = time.perf_counter()
    ev[0].record(stream)
    for i in range(K):
        step()
        ev[i + 1].record(stream)
    stream.sync()
    gpu.synchronize()
    wall = time.perf_counter() - t_wall
    per = np.array([ev[i].elapsed_ms(ev[i + 1]) for i in range(K)])
    total_ms = ev[0].elapsed_ms(ev[K])
    ms_step = total_ms / K
    value = mesh.nnz * K / (total_ms * 1e-3)
    B = algorithmic_bytes(mesh.nbElem, mesh.nbNodes, mesh.nnz, op)
    # one fused launch per step on the tiled schedule: mean launch time = mean step time
    kern_s = float(per.mean()) * 1e-3
    achieved = B / kern_s / 1e9
    kname = f"k_tiled<{args.op},{args.precision}>" if plan is not None else "colored step"
    # HBM traffic per launch is not counted in this run: it is the committed rocprofv3
    # --pmc summary of the same kernel on the same workload (FETCH_SIZE + WRITE_SIZE per
    # dispatch, scaled by the copy-kernel calibration the summary states), labelled so
    traffic = traffic_src = None
    size = f"delaunay{args.delaunay}" if args.delaunay else str(args.n)
    tname = f"traffic_{args.op}_{size}_{args.schedule}_{args.precision}.json"
    tpath = os.path.join(ROOT, "profiles", tname)
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        traffic = tj.get("bytes_per_launch")
        traffic_src = {"source": f"profiles/{tname} (rocprofv3 --pmc, not measured in this run)",
                       "measured": tj.get("measured"), "fetch_bytes": tj.get("fetch_bytes"),
                       "write_bytes": tj.get("write_bytes")}
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_from": traffic_src,
                "kernel": kname, "algorithmic_bytes_per_launch": B,
                "mean_launch_us": round(kern_s * 1e6, 2)}
    if floor is not None:
        floor["kernel_over_tile_rows"] = round(kern_s * 1e6 / floor["tile_rows_us"], 3)
        roofline["output_write_floor"] = floor
    log(f"GPU: {ms_step * 1e3:.1f} us/step, min {per.min() * 1e3:.1f} us, "
        f"{achieved:.0f} GB/s algorithmic ({100 * achieved / HBM_PEAK_GBS:.1f}% of peak)")

    # the timed precision's results (the last step's output), checked below
    res_v, res_p = dm.values.numpy(), dm.prec.numpy()
    cpu = cpu_extra = accuracy = None
    if not args.no_cpu:
        cpu, cpu_extra, (ov, opv) = cpu_baselines(mesh, op, value, args)
        # the GPU results against the oracle's FP64 iteration just timed (COLORING order)
        accuracy = {"vs": "oracle FP64 iteration, COLORING order, same mesh (the CPU leg)",
                    "values_bit_exact": bool(np.array_equal(res_v, ov)),
                    "prec_bit_exact": bool(np.array_equal(res_p, opv)),
                    "values_rel_fro": rel_fro(res_v, ov), "prec_rel_fro": rel_fro(res_p, opv)}
        del ov, opv
    elif prec_mode != minifem.PREC_FP64:
        # against one FP64 step of the same kernel (bit-identical to the oracle's COLORING
        # order, a property the GPU tests hold at this size); a plan sized for the float
        # stage runs the mixed modes only, so the FP64 step gets a plan of its own
        if plan is not None and plan.stage_precision != minifem.PREC_FP64:
            p64 = gpu.Plan(mesh, **dict(params, stage_precision=minifem.PREC_FP64))
            p64.assemble(dm.coord, dm.coef, dm.values, dm.prec, precision=minifem.PREC_FP64,
                         stream=stream)
            del p64
        else:
            step(minifem.PREC_FP64)
        stream.sync()
        v64, p64 = dm.values.numpy(), dm.prec.numpy()
        accuracy = {"vs": "FP64 step of the same schedule (bit-identical to the oracle)",
                    "values_rel_fro": rel_fro(res_v, v64), "prec_rel_fro": rel_fro(res_p, p64)}
        del v64, p64
    if accuracy is not None:
        log(f"accuracy: {accuracy}")

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "CSR blocks/s", "n_gpus": 1,
        "steps": K, "warmup": args.warmup, "ms_per_step": round(ms_step, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec_mode == minifem.PREC_FP64 else "f32 products, f64 accumulate",
        "data": "synthetic",
        "config": {"workload": f"{args.op} {shape}: "
                               f"{mesh.nbNodes} nodes, {mesh.nbElem} tets, {mesh.nnz} CSR blocks; "
                               "assembly + block-Jacobi prec per step",
                   "schedule": args.schedule, "precision": args.precision,
                   "tile_params": params if plan is not None else None,
                   "wall_s": round(wall, 4)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "cpu_baselines": cpu_extra,
        "accuracy": accuracy,
        "tuning": tuning,
    }
    if (not args.no_layers and not args.delaunay and args.n == 100 and args.op == "ela" and
            plan is not None):
        # the consumer of the step's outputs, on the operator and prec just assembled
        if prec_mode == minifem.PREC_FP64:
            step()
            stream.sync()
            out["pcg"] = pcg_forms(mesh, dm, plan, op, 100, 5)
        # the other two workloads of the headline metric, on the same GPU in the same run
        del plan, dm
        out["batched"] = batched_leg(args)
        out["served"] = served_leg(args)
    print(json.dumps(out), flush=True)


def host_cpu():
    """Model name and the cores this process may run on (this host's CPU share)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "cpus_visible": os.cpu_count(),
            "cpus_allowed": len(os.sched_getaffinity(0))}


def cpu_baselines(mesh, op, gpu_value, args):
    """The host-core baselines on the workload mesh, as the reference times them
    (src/FEM.cc:177-257: per-phase averages, iteration 0 excluded when there are more than
    one, src/FEM.cc:182): the oracle's REF order (one thread) and its OpenMP COLORING
    build twice, as written (serial element loop within each color, src/assembly.cc:362;
    parallel zero-fill and prec loops) and with each color's elements in parallel. The
    GPU speedup is quoted against the faster COLORING run (cpu_baseline) and against REF.
    Returns (cpu_baseline, the three runs, (values, prec) of the per-color parallel run)."""
    import numpy as np
    import pyoracle
    runs = {}
    keep = None
    threads = pyoracle.num_threads()
    for name, mode, iters, cores in (("coloring_omp_parallel_colors", 2, args.cpu_iters, threads),
                                     ("coloring_omp_as_written", 1, 2, threads),
                                     ("ref", 0, 2, 1)):
        t0 = time.perf_counter()
        v, p, tus = pyoracle.fem_loop(mesh.coord, mesh.row, mesh.col, mesh.elemToNode,
                                      mesh.checkBounds, mesh.colorToElem, op, mesh.nbNodes,
                                      iters, mode, coef=mesh.coef)
        s_it = float(np.sum(tus)) * 1e-6
        runs[name] = {"value": round(mesh.nnz / s_it, 1), "unit": "CSR blocks/s", "cores": cores,
                      "kind": "port", "s_per_iteration": round(s_it, 4),
                      "phase_us": [round(float(x), 1) for x in tus],
                      "iterations": iters, "timed_iterations": max(1, iters - 1),
                      "gpu_speedup": round(gpu_value / (mesh.nnz / s_it), 1)}
        log(f"CPU {name}: {s_it:.3f} s/iteration ({time.perf_counter() - t0:.1f}s, {cores} threads)")
        if mode == 2:
            keep = (v, p)
        del v, p
    best = min(("coloring_omp_parallel_colors", "coloring_omp_as_written"),
               key=lambda k: runs[k]["s_per_iteration"])
    b = runs[best]
    cpu = {"value": b["value"], "unit": "CSR blocks/s", "cores": b["cores"], "kind": "port",
           "sample": f"the oracle's OpenMP COLORING iteration ({best}, {b['cores']} threads, "
                     f"-O2), {b['iterations']} iterations of the same "
                     f"{f'{args.delaunay}-point Delaunay' if args.delaunay else f'{args.n}^3'} {args.op} "
                     f"mesh, iteration 0 excluded: {b['s_per_iteration']} s per iteration "
                     f"(assembly + prec_init + prec_inversion)",
           "gpu_speedup": b["gpu_speedup"], "gpu_speedup_vs_ref_1_core": runs["ref"]["gpu_speedup"],
           "host": host_cpu()}
    return cpu, runs, keep


def batched_leg(args):
    """The batched workload in the default line: 64 meshes (5k-200k nodes, lap and ela
    alternating) in the union form with the library's default schedule, one captured
    graph per batch, FP64 and mixed; every mixed mesh's error against FP64."""
    import numpy as np

    import minifem
    from minifem.batched import Batch, batch_meshes
    t0 = time.time()
    meshes = batch_meshes()
    B = sum(algorithmic_bytes(m.nbElem, m.nbNodes, m.nnz, m.operatorID) for m in meshes)
    out = {"workload": f"{len(meshes)} jittered cubes of 5k-200k nodes, lap/ela alternating: "
                       f"{sum(m.nbNodes for m in meshes)} nodes, "
                       f"{sum(m.nnz for m in meshes)} CSR blocks",
           "form": "union (one plan per operator), captured in one HIP graph; the mixed "
                   "batch's tiles sized by its float stage"}
    ref = None
    for name, prec in (("fp64", minifem.PREC_FP64), ("mixed", minifem.PREC_MIXED)):
        u = Batch(meshes, precision=prec, union=True).capture()
        out["launches_per_batch"] = u.launches()
        ms = u.time_ms(iters=max(5, min(args.steps, 20)), warmup=args.warmup)
        gbs = B / (ms * 1e-3) / 1e9
        out[name] = {"ms_per_batch": round(ms, 4), "blocks_per_s": round(u.blocks() / (ms * 1e-3), 1),
                     "algorithmic_GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
        got = u.results()
        if ref is None:
            ref = got
        else:
            out[name]["max_rel_err_vs_fp64"] = {
                "values": max(rel_fro(v, r[0]) for (v, _), r in zip(got, ref)),
                "prec": max(rel_fro(p, r[1]) for (_, p), r in zip(got, ref))}
        del got, u
    out["mixed_vs_fp64_throughput"] = round(out["fp64"]["ms_per_batch"] / out["mixed"]["ms_per_batch"], 3)
    log(f"batched: {out} ({time.time() - t0:.1f}s)")
    del ref
    return out


def make_server(meshes, budget, precision, slots):
    """The served forms: slots=1 is the plain server (one request per launch, one launch at
    a time); 4 slots is continuous batching (4 streams, requests that arrived together
    served in pairs by one launch)."""
    from minifem.serve import Server
    return Server(meshes, budget, slots=slots, precision=precision,
                  max_group=1 if slots == 1 else 2)


def served_leg(args, rate=5000.0, count=600, saturating=100000.0):
    """The served workload in the default line: 16 resident meshes (1k-885k nodes), a
    seeded Poisson stream under an HBM budget of half the footprints' sum, served by
    plain FP64 one request at a time and by continuous batching over 4 slots in FP64 and
    in mixed precision. Each server takes the stream at `rate` (latency) and the same
    requests at a saturating rate (throughput: every form busy all the time). Latency
    runs from arrival to the end of the request's last kernel (the results stay in HBM for
    the consumer; readback is not included)."""
    import numpy as np

    import minifem
    from minifem.serve import make_requests, retime, served_meshes
    t0 = time.time()
    meshes = served_meshes()
    reqs = make_requests(meshes, count, rate)
    fps = [8 * m.dim * (m.nnz + m.nbNodes) + 8 * max(m.coord.size, m.coef.size) for m in meshes]
    budget = max(int(0.5 * sum(fps)), int(1.05 * max(fps)) + 4096)
    out = {"workload": f"{len(meshes)} resident meshes of {min(m.nbNodes for m in meshes)}-"
                       f"{max(m.nbNodes for m in meshes)} nodes, {count} Poisson requests at "
                       f"{rate:g}/s (coefficient field or coordinates), HBM budget "
                       f"{budget / 1e9:.2f} GB of {sum(fps) / 1e9:.2f} GB footprints",
           "latency": "arrival to the end of the request's last kernel; readback excluded"}
    sample = reqs[:8]
    ref = {}
    for name, prec, slots in (("plain_fp64_1_slot", minifem.PREC_FP64, 1),
                              ("batched_fp64_4_slots", minifem.PREC_FP64, 4),
                              ("batched_mixed_4_slots", minifem.PREC_MIXED, 4)):
        srv = make_server(meshes, budget, prec, slots)
        srv.stage(reqs)
        st = srv.run(reqs)
        out[name] = {k: round(v, 4) for k, v in st.items()}
        sat = srv.run(retime(reqs, saturating))
        out[name]["saturated"] = {k: round(v, 4) for k, v in sat.items()
                                  if k in ("p50_ms", "p99_ms", "requests_per_s", "blocks_per_s")}
        got = {r.rid: srv.run_alone(r) for r in sample}
        if not ref:
            ref = got
        else:
            out[name]["sample_max_rel_err_vs_plain_fp64"] = {
                "requests": len(sample),
                "values": max(rel_fro(got[k][0], ref[k][0]) for k in got),
                "prec": max(rel_fro(got[k][1], ref[k][1]) for k in got)}
        del got
        if name != "batched_mixed_4_slots":  # the last server also admits a new mesh
            del srv
    # a topology that was not resident: its setup (coloring, permutation, CSR, mask) on the
    # GPU, then the plan, device arrays and a warm launch, while the server holds its 16
    # meshes; the same setup on the host is timed beside it
    from minifem.setup import Mesh, generate_coef, generate_cube
    coord, e2n, code, nnz = generate_cube(100, 0.2, 4242, 4243)
    coef = generate_coef(1, e2n.size // 4, 4244)
    t0h = time.perf_counter()
    hmesh = Mesh(coord, e2n, code, nnz, 1, coef=coef)
    t1 = time.perf_counter()
    mesh = Mesh(coord, e2n, code, nnz, 1, coef=coef, setup="device")
    t2 = time.perf_counter()
    same = all(np.array_equal(getattr(hmesh, a), getattr(mesh, a))
               for a in ("elemToNode", "coef", "colorPerm", "colorToElem", "row", "col"))
    del hmesh
    idx = srv.add_mesh(mesh)
    t3 = time.perf_counter()
    out["new_mesh_admission"] = {"mesh": f"ela 100^3 node-shuffled, {mesh.nbNodes} nodes, "
                                         f"{mesh.nbElem} tets",
                                 "device_setup_s": round(t2 - t1, 3),
                                 "host_setup_s": round(t1 - t0h, 3),
                                 "device_setup_equals_host": bool(same),
                                 "plan_upload_warm_s": round(t3 - t2, 3),
                                 "total_s": round(t3 - t1, 3), "mesh_index": idx}
    del srv, mesh
    rps = {k: out[k]["saturated"]["requests_per_s"]
           for k in ("plain_fp64_1_slot", "batched_fp64_4_slots", "batched_mixed_4_slots")}
    out["saturated_rate_hz"] = saturating
    out["throughput_ratios"] = {  # at the saturating rate
        "batched_fp64_vs_plain": round(rps["batched_fp64_4_slots"] / rps["plain_fp64_1_slot"], 3),
        "batched_mixed_vs_plain": round(rps["batched_mixed_4_slots"] / rps["plain_fp64_1_slot"], 3),
        "batched_mixed_vs_batched_fp64_same_slots":
            round(rps["batched_mixed_4_slots"] / rps["batched_fp64_4_slots"], 3)}
    log(f"served: {out} ({time.time() - t0:.1f}s)")
    return out


def rel_fro(a, b):
    """Relative Frobenius error of a against b over b's finite entries."""
    import numpy as np
    m = np.isfinite(b)
    n = float(np.linalg.norm(b[m]))
    d = float(np.linalg.norm(a[m] - b[m]))
    return d / n if n > 0 else d


def pcg_iteration_bytes(N, nnz, op):
    """HBM bytes one PCG iteration of mf_pcg moves on n = N d dofs, each array once: the
    operator product (values, col_idx, row_ptr, the per-node mask byte, the masked
    direction it gathers, the direction and the product), the fused update and
    block-Jacobi step (x, p, q, r and prec in; x, r and z out) and the direction update
    (z and p in, p and its masked copy out, the mask)."""
    d = 3 if op == 1 else 1
    n = N * d
    spmv = 8 * d * d * nnz + 4 * nnz + 4 * (N + 1) + N + 24 * n
    return {"spmv": spmv, "total": spmv + (56 * n + 8 * d * d * N) + 32 * n + N}


def pcg_forms(mesh, dm, plan, op, iters, warmup):
    """Block-Jacobi PCG on the operator and prec in dm (rtol 0: every iteration runs),
    right-hand side P 1, x0 = 0, in the mesh's own node numbering (mf_pcg) and in the
    plan's tile order (mf_pcg_ordered). Per iteration: (time of `iters` iterations - time of
    0) / iters; the call's fixed part (mask, first residual, and for the ordered form the
    renumbered copy of the system) is the time of 0."""
    import numpy as np

    import minifem
    import minifem.gpu as gpu
    d = 3 if op == minifem.OP_ELA else 1
    order = plan.row_order()
    b = gpu.DeviceArray.from_numpy(np.ones(mesh.nbNodes * d))
    x = gpu.DeviceArray.zeros(mesh.nbNodes * d, np.float64)
    B = pcg_iteration_bytes(mesh.nbNodes, mesh.nnz, op)

    def run(k, o):
        x.fill(0.0)
        gpu.synchronize()
        t = time.perf_counter()
        it, rel = gpu.pcg(op, mesh.nbNodes, dm.row, dm.col, dm.values, dm.prec, dm.checkBounds,
                          b, x, rtol=1e-300, max_iter=k, order=o)
        return time.perf_counter() - t, it, rel

    forms = {}
    for name, o in (("mesh_order", None), ("plan_order", order)):
        run(max(1, warmup), o)
        t_fixed = min(run(0, o)[0] for _ in range(3))
        t_k, it, rel = min(run(iters, o) for _ in range(2))
        us = (t_k - t_fixed) / max(it, 1) * 1e6
        forms[name] = {"us_per_iteration": round(us, 2), "fixed_ms": round(t_fixed * 1e3, 3),
                       "iterations": it, "rel_residual": rel,
                       "iteration_GBs": round(B["total"] / us / 1e3, 1),
                       "frac_of_hbm_peak": round(B["total"] / us / 1e3 / HBM_PEAK_GBS, 4)}
        log(f"pcg {name}: {forms[name]}")
    return {"consumer": "block-Jacobi PCG on the step's operator and prec (mf_pcg, "
                        "mf_pcg_ordered with the plan's row order)",
            "rel_residual": "||P 1 - K x|| / ||P 1|| after the timed iterations; CG minimizes "
                            "the error's energy norm, and the residual's 2-norm first grows (on "
                            "the 20^3 ela cube, 16x at iteration 5 before it falls to 1e-9 at "
                            "400); the 100^3 operator is far worse conditioned (a tet with "
                            "|det| = 1.5e-5 h^3, entries scaling as 1/det^2): 0.11 after 20,000 "
                            "iterations, so only the time per iteration is measured here",
            "iteration_bytes": B, "forms": forms}


def bench_pcg(args):
    """--pcg K: the PCG consumer alone on the --op/--n bench cube (pcg_forms)."""
    import minifem
    import minifem.gpu as gpu
    gpu.init(0)
    op = minifem.OP_ELA if args.op == "ela" else minifem.OP_LAP
    t0 = time.time()
    coord, e2n, code, nnz = minifem.generate_cube(args.n, 0.25, 3, 203)
    mesh = minifem.Mesh(coord, e2n, code, nnz, op, coef=minifem.generate_coef(op, e2n.size // 4, 103))
    dm = gpu.DeviceMesh(mesh)
    plan = gpu.Plan(mesh)
    plan.assemble(dm.coord, dm.coef, dm.values, dm.prec)
    gpu.synchronize()
    log(f"pcg: {args.op} {args.n}^3 assembled ({time.time() - t0:.1f}s)")
    res = pcg_forms(mesh, dm, plan, op, args.pcg, args.warmup)
    best = res["forms"]["plan_order"]
    d = 3 if op == minifem.OP_ELA else 1
    out = {"metric": "PCG iterations/s on the assembled operator", "unit": "iterations/s",
           "value": round(1e6 / best["us_per_iteration"], 1), "form": "plan_order",
           "workload": f"{args.op} {args.n}^3 node-shuffled: {mesh.nbNodes * d} dofs, "
                       f"{mesh.nnz} CSR blocks, {args.pcg} iterations", **res}
    print(json.dumps(out), flush=True)


def bench_batched(args):
    """64 independent meshes assembled by one replay of the captured batch graph, in the
    per-mesh form (one launch per mesh on 8 side streams) and the union form (one
    persistent launch per operator over the disjoint union of its meshes)."""
    import numpy as np

    import minifem
    import minifem.gpu as gpu
    from minifem.batched import Batch, batch_meshes, tune_batch
    gpu.init(0)
    t0 = time.time()
    meshes = batch_meshes()
    log(f"batch: {len(meshes)} meshes, {sum(m.nbNodes for m in meshes)} nodes, "
        f"{sum(m.nnz for m in meshes)} blocks ({time.time() - t0:.1f}s)")
    out = {"metric": "batched CSR blocks assembled/s (64 meshes, one graph launch)",
           "unit": "CSR blocks/s", "n_gpus": 1, "data": "synthetic"}
    B = sum(algorithmic_bytes(m.nbElem, m.nbNodes, m.nnz, m.operatorID) for m in meshes)
    fp64_results = {}
    forms = [f for f in args.batch_forms.split(",") if f]
    if not forms or any(f not in ("per_mesh", "union") for f in forms):
        raise SystemExit(f"--batch-forms: per_mesh and/or union, got {args.batch_forms!r}")
    fixed = None
    if args.batch_params:
        lds, blk, rows = (int(x) for x in args.batch_params.split(","))
        fixed = dict(tile_lds_bytes=lds, block_threads=blk, max_tile_rows=rows)
    for form, union in (("per_mesh", False), ("union", True)):
        if form not in forms:
            continue
        if fixed is None:
            best, res = tune_batch(meshes, log=log, union=union)
        else:
            best = fixed
        out[form] = {"schedule": best}
        for name, prec in (("fp64", minifem.PREC_FP64), ("mixed", minifem.PREC_MIXED)):
            b = Batch(meshes, precision=prec, union=union, **best).capture()
            ms_graph = b.time_ms(iters=args.steps, warmup=args.warmup)
            ms_eager = b.time_ms(iters=args.steps, warmup=args.warmup, graph=False)
            out[form][name] = {"ms_per_batch": round(ms_graph, 4),
                               "ms_per_batch_eager": round(ms_eager, 4),
                               "launches": b.launches(),
                               "value": round(b.blocks() / (ms_graph * 1e-3), 1)}
            got = b.results()
            if name == "fp64":
                fp64_results[form] = got
            else:  # the error of every mixed result against the FP64 one, worst mesh
                ref = fp64_results[form]
                out[form][name]["max_rel_err_vs_fp64"] = {
                    "values": max(rel_fro(v, r[0]) for (v, _), r in zip(got, ref)),
                    "prec": max(rel_fro(p, r[1]) for (_, p), r in zip(got, ref))}
            log(f"{form} {name}: {out[form][name]}")
            del b, got
        gbs = B / (out[form]["fp64"]["ms_per_batch"] * 1e-3) / 1e9
        out[form]["roofline_fp64"] = {"achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                                      "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4)}
    if len(fp64_results) == 2:
        out["union_equals_per_mesh_fp64"] = all(
            np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
            for a, b in zip(fp64_results["per_mesh"], fp64_results["union"]))
    if len(fp64_results) == 2:
        out["union_vs_per_mesh_fp64"] = round(out["union"]["fp64"]["value"] /
                                              out["per_mesh"]["fp64"]["value"], 3)
    out["value"] = max(out[f]["fp64"]["value"] for f in fp64_results)
    print(json.dumps(out), flush=True)


def bench_serve(args):
    """Continuous batching on 16 resident meshes against plain FP64 one request at a time."""
    import numpy as np

    import minifem
    import minifem.gpu as gpu
    from minifem.serve import make_requests, served_meshes
    gpu.init(0)
    t0 = time.time()
    meshes = served_meshes()
    reqs = make_requests(meshes, args.requests, args.rate)
    fps = [8 * m.dim * (m.nnz + m.nbNodes) + 8 * max(m.coord.size, m.coef.size) for m in meshes]
    budget = max(int(0.5 * sum(fps)), int(1.05 * max(fps)) + 4096)
    log(f"serve: {len(meshes)} meshes up to {max(m.nbNodes for m in meshes)} nodes, "
        f"{len(reqs)} requests at {args.rate}/s, HBM budget {budget / 1e9:.2f} GB "
        f"({time.time() - t0:.1f}s)")
    out = {"metric": "served assembly requests: p50/p99 latency and throughput vs plain FP64",
           "unit": "ms", "n_gpus": 1, "data": "synthetic", "requests": len(reqs),
           "rate_hz": args.rate, "hbm_budget_bytes": budget, "higher_is_better": False}
    sample = reqs[:8]  # requests whose results are checked against plain FP64
    ref = {}
    for name, prec, slots in (("plain_fp64_one_at_a_time", minifem.PREC_FP64, 1),
                              ("batched_fp64", minifem.PREC_FP64, 4),
                              ("batched_mixed", minifem.PREC_MIXED, 4)):
        srv = make_server(meshes, budget, prec, slots)
        srv.stage(reqs)
        st = srv.run(reqs)
        out[name] = {k: round(v, 4) for k, v in st.items()}
        got = {r.rid: srv.run_alone(r) for r in sample}
        if not ref:
            ref = got
        else:
            out[name]["sample_max_rel_err_vs_plain_fp64"] = {
                "requests": len(sample),
                "values": max(rel_fro(got[k][0], ref[k][0]) for k in got),
                "prec": max(rel_fro(got[k][1], ref[k][1]) for k in got)}
        log(f"{name}: {out[name]}")
        del srv, got
    out["throughput_vs_plain"] = {
        k: round(out[k]["requests_per_s"] / out["plain_fp64_one_at_a_time"]["requests_per_s"], 3)
        for k in ("batched_fp64", "batched_mixed")}
    out["value"] = out["batched_fp64"]["p99_ms"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
