"""Headline benchmark: the FP64 assembly-free EbE matvec on one MI355X.

Workload (the GPU drop-in at real shape): 3D linear elasticity on the unit
cube, 64^3 quadratic 20-node serendipity hexahedra, 3 DoFs per node
(262,144 elements, 3,257,475 DoFs, 60 x 60 element matrices), E = 1,
nu = 0.3, 3x3x3 Gauss points, K_e scaled by E_e ~ U[0.5, 2) from a seed,
u ~ U[-1, 1), penalty BC (diagonal 1e8) on the z = 0 face, deterministic
transposed-gather scatter.  A step is one operator application v = A u with
every input already resident in HBM.

Prints one JSON line: GDoF/s (whole job), ms per step, the roofline of the
dominant kernel (the element kernel: algorithmic bytes / its mean duration,
hipEvents on its launch stream) and a CPU baseline (the oracle's restatement
of the reference OpenMP matvec_ghosted_OMP, timed on this host on a bounded
sample of the same elements).
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "EbE matvec GDoF/s + %HBM peak; served CG solves/s, p99 latency @1e-10 rel res"
HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(p, n_sample_layers: int, reps: int, threads: int) -> dict:
    """The oracle's matvec_ghosted_OMP (default build: row-major scalar GEMV +
    omp atomic) on the first n_sample_layers element layers of the workload,
    scaled to the whole mesh (time per element x all elements)."""
    import numpy as np
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as orc
    E = p.n_elem
    per_layer = E // p.mesh.dims[2]
    ns = min(E, n_sample_layers * per_layer)
    gm = p.gmap[:ns]
    K = p.element_matrices(0, ns)
    u = p.random_u()
    orc.matvec(gm, K, u, n_total=p.n_dofs, nthreads=threads)  # warm-up
    t0 = time.perf_counter()
    for _ in range(reps):
        orc.matvec(gm, K, u, n_total=p.n_dofs, nthreads=threads)
    dt = (time.perf_counter() - t0) / reps
    t_full = dt * E / ns
    return {"value": p.n_dofs / t_full / 1e9, "unit": "GDoF/s", "cores": threads, "kind": "port",
            "sample": f"{ns} of {E} elements ({n_sample_layers} element layers), {reps} matvecs, "
                      f"{dt * 1e3:.1f} ms each, scaled by element count; oracle matvec_ghosted_OMP "
                      f"(row-major scalar GEMV + omp atomic), {threads} OpenMP threads"}


def load_traffic(name: str):
    """HBM bytes per launch of the element kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (None when absent)."""
    f = ROOT / "profiles" / f"{name}_traffic.md"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith("bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=64, help="elements per cube edge")
    ap.add_argument("--scatter", choices=["deterministic", "atomic"], default="deterministic")
    ap.add_argument("--cpu-layers", type=int, default=4, help="element layers in the CPU baseline sample")
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after the timed region, copy 1 GiB device-to-device 5 times (known bytes for the "
                         "FETCH_SIZE / WRITE_SIZE counter scale)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU's share; --gpus {args.gpus} is not run here (multi-GPU runs are out of "
              "scope for this build)", file=sys.stderr)
        return 2

    import numpy as np
    import torch
    from amat_amd import BCOP, PREC, SCATTER, Maps, aMatFree, problems

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    t0 = time.time()
    p = problems.elasticity_hex20(args.n)
    log(f"[bench] mesh hex20 {args.n}^3: {p.n_elem} elements, {p.n_dofs} DoFs ({time.time() - t0:.1f}s)")
    scatter = SCATTER.DETERMINISTIC if args.scatter == "deterministic" else SCATTER.ATOMIC
    op = p.build_operator(aMatFree, Maps, precision=PREC.FP64, scatter=scatter, bcop=BCOP.PENALTY_DIAG, penalty=1e8)
    info = op.info()
    log(f"[bench] operator finalized: {info.device_bytes / 1e9:.2f} GB on device ({time.time() - t0:.1f}s)")
    s = torch.cuda.current_stream()
    u = torch.from_numpy(p.random_u()).cuda()
    v = torch.empty_like(u)
    for _ in range(args.warmup):
        op.matmult_dev(v, u, s)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(s)
    for _ in range(args.steps):
        op.matmult_dev(v, u, s)
    ev1.record(s)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / args.steps
    log(f"[bench] {args.steps} steps: {ms:.4f} ms/step")
    # dominant kernel: mean duration of the element-kernel launches alone
    ms_elem, ms_gather, ms_tot = op.time_kernels_dev(args.steps, v, u, s)
    log(f"[bench] kernel timers: element {ms_elem:.4f} ms, gather {ms_gather:.4f} ms")
    if args.calibrate_copy:
        src = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    E, n, N = p.n_elem, p.dofs_per_elem, p.n_dofs
    elem_bytes = E * (n * n * 8 + 4 * n) + N * 8       # K + map + u read
    alg_bytes = info.alg_bytes_fp64                     # + v write: the whole matvec
    achieved = elem_bytes / (ms_elem * 1e-3)
    traffic = load_traffic(f"hex20_{args.n}_fp64_ebe")  # rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated
    out = {
        "metric": METRIC,
        "value": N / (ms * 1e-3) / 1e9,
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"hex20 elasticity {args.n}^3 FP64 EbE matvec, E_e~U[0.5,2), penalty 1e8 on z=0, "
                               f"{args.scatter} scatter", "elements": E, "dofs": N, "dofs_per_element": n,
                   "parallelism": "1 GPU (one rank's share)"},
        "matvec_hbm_frac": alg_bytes / (ms * 1e-3) / HBM_PEAK,
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic, "kernel": "ebe_tile_kernel<60,double>",
                     "alg_bytes_per_launch": elem_bytes, "ms_per_launch": ms_elem, "gather_ms": ms_gather},
    }
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log(f"[bench] CPU baseline on {threads} threads ...")
        out["cpu_baseline"] = cpu_baseline(p, args.cpu_layers, args.cpu_reps, threads)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
