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
8 TB/s HBM3E peak), and the host-core baseline (the CPU oracle's REF-order iteration on
the same mesh, timed here).
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
    # schedule of the tiled plan, as tuned for the 100^3 ela mesh on MI355X (the library's
    # own defaults suit smaller meshes; minifem.tune picks per mesh)
    ap.add_argument("--tile-lds", type=int, default=150000)
    ap.add_argument("--block", type=int, default=1024)
    ap.add_argument("--max-rows", type=int, default=254)
    ap.add_argument("--pair-window", type=int, default=1 << 30)
    ap.add_argument("--copy", type=int, default=0,
                    help="instead of the FEM step, time the plain HBM copy kernel on this many "
                         "bytes (bandwidth reference and byte-counter calibration)")
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

    op = minifem.OP_ELA if args.op == "ela" else minifem.OP_LAP
    prec_mode = minifem.PREC_FP64 if args.precision == "fp64" else minifem.PREC_MIXED
    gpu.init(0)
    t0 = time.time()
    coord, e2n, code, nnz = minifem.generate_cube(args.n, 0.25, 3, 203)
    coef = minifem.generate_coef(op, e2n.size // 4, 103)
    mesh = minifem.Mesh(coord, e2n, code, nnz, op, coef=coef)
    log(f"setup: {mesh.nbNodes} nodes, {mesh.nbElem} tets, {mesh.nnz} blocks, "
        f"{mesh.nbColors} colors ({time.time() - t0:.1f}s)")
    dm = gpu.DeviceMesh(mesh)
    plan = None
    if args.schedule == "tiled":
        t1 = time.time()
        plan = gpu.Plan(mesh, tile_lds_bytes=args.tile_lds, block_threads=args.block,
                        max_tile_rows=args.max_rows, pair_window=args.pair_window)
        log(f"plan: {plan.stats()} ({time.time() - t1:.1f}s)")
    stream = gpu.Stream()

    def step():
        if plan is not None:
            plan.assemble(dm.coord, dm.coef, dm.values, dm.prec, precision=prec_mode,
                          stream=stream)
        else:
            dm.iterate_colored(precision=prec_mode, stream=stream)

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
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{args.op}_{args.n}_{args.schedule}_"
                         f"{args.precision}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": kname, "algorithmic_bytes_per_launch": B,
                "mean_launch_us": round(kern_s * 1e6, 2)}
    log(f"GPU: {ms_step * 1e3:.1f} us/step, min {per.min() * 1e3:.1f} us, "
        f"{achieved:.0f} GB/s algorithmic ({100 * achieved / HBM_PEAK_GBS:.1f}% of peak)")

    cpu = None
    if not args.no_cpu:
        import pyoracle
        t2 = time.perf_counter()
        _, _, tus = pyoracle.fem_loop(mesh.coord, mesh.row, mesh.col, mesh.elemToNode,
                                      mesh.checkBounds, mesh.colorToElem, op, mesh.nbNodes, 1, 0,
                                      coef=mesh.coef)
        cpu_s = float(np.sum(tus)) * 1e-6
        log(f"CPU oracle REF: {cpu_s:.2f} s/iteration ({time.perf_counter() - t2:.1f}s)")
        cpu = {"value": round(mesh.nnz / cpu_s, 1), "unit": "CSR blocks/s", "cores": 1,
               "kind": "port",
               "sample": f"1 full iteration (assembly + prec_init + prec_inversion, REF order, "
                         f"serial, -O2) of the same {args.n}^3 {args.op} mesh: "
                         f"{cpu_s:.2f} s; phases us {[round(float(x), 1) for x in tus]}",
               "gpu_speedup": round(value / (mesh.nnz / cpu_s), 1)}

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "CSR blocks/s", "n_gpus": 1,
        "steps": K, "warmup": args.warmup, "ms_per_step": round(ms_step, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec_mode == minifem.PREC_FP64 else "f32 products, f64 accumulate",
        "data": "synthetic",
        "config": {"workload": f"{args.op} {args.n}^3 jittered Kuhn cube, node-shuffled: "
                               f"{mesh.nbNodes} nodes, {mesh.nbElem} tets, {mesh.nnz} CSR blocks; "
                               "assembly + block-Jacobi prec per step",
                   "schedule": args.schedule, "precision": args.precision,
                   "wall_s": round(wall, 4)},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
