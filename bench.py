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


def cpu_baseline(p, n_sample_layers: int, seconds: float, threads: int) -> dict:
    """The oracle's matvec_ghosted_OMP (default build: row-major scalar GEMV +
    omp atomic) on the first n_sample_layers element layers of the workload
    (all of them by default), repeated for about `seconds` of CPU time and
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

    def timed(layout, library=None):
        t0 = time.perf_counter()
        orc.matvec(gm, K, u, n_total=p.n_dofs, nthreads=threads, layout=layout, library=library)  # warm-up
        t1 = time.perf_counter() - t0
        reps = max(1, int(seconds / max(t1, 1e-6)))
        t0 = time.perf_counter()
        for _ in range(reps):
            orc.matvec(gm, K, u, n_total=p.n_dofs, nthreads=threads, layout=layout, library=library)
        return (time.perf_counter() - t0) / reps, reps

    dt, reps = timed(0)
    # the VECTORIZED_OPENMP variant (column AXPY with omp simd): K_e is symmetric, so
    # its column-major storage is the same array
    dt1, reps1 = timed(1)
    # the same loop built for this host with -O3 -march=native, the flags the
    # reference's AVX builds add (CMakeLists.txt:93-104): the "OpenMP/AVX" number
    import tempfile
    with tempfile.TemporaryDirectory() as tmp:
        nat = orc.native_library(tmp)
        dt2, reps2 = timed(1, nat)
    rates = {"default": p.n_dofs / (dt * E / ns) / 1e9, "vectorized_openmp": p.n_dofs / (dt1 * E / ns) / 1e9,
             "vectorized_native": p.n_dofs / (dt2 * E / ns) / 1e9}
    best = max(rates, key=rates.get)
    t_best = {"default": dt, "vectorized_openmp": dt1, "vectorized_native": dt2}[best]
    return {"value": rates[best], "unit": "GDoF/s", "cores": threads, "kind": "port",
            "sample": f"{ns} of {E} elements ({ns // per_layer} element layers), about {seconds:.0f} s of matvecs "
                      f"per variant, scaled by element count; oracle matvec_ghosted_OMP, {threads} OpenMP threads; "
                      f"value = the fastest variant ({best}, {t_best * 1e3:.1f} ms per matvec)",
            "cpu_model": cpu_model(),
            "variants_GDoF_s": rates,
            "variants": {"default": f"row-major scalar GEMV + omp atomic, -O2 (the reference's default build), "
                                    f"{reps} matvecs, {dt * 1e3:.1f} ms each",
                         "vectorized_openmp": f"column AXPY + omp simd (VECTORIZED_OPENMP), -O2, {reps1} matvecs, "
                                              f"{dt1 * 1e3:.1f} ms each",
                         "vectorized_native": f"column AXPY + omp simd, -O3 -march=native (the AVX builds' flags), "
                                              f"{reps2} matvecs, {dt2 * 1e3:.1f} ms each"}}


def cpu_model() -> str:
    """The host CPU: model name, family/model/stepping and its widest SIMD."""
    info = {}
    try:
        for line in open("/proc/cpuinfo"):
            if ":" in line:
                k, v = (x.strip() for x in line.split(":", 1))
                if k in ("model name", "cpu family", "model", "stepping", "flags") and k not in info:
                    info[k] = v
            if not line.strip() and info:
                break
    except OSError:
        return ""
    flags = set(info.get("flags", "").split())
    simd = next((f for f in ("avx512f", "avx2", "avx", "sse4_2") if f in flags), "sse2")
    return (f"{info.get('model name', '?')} (family {info.get('cpu family', '?')} model {info.get('model', '?')} "
            f"stepping {info.get('stepping', '?')}, {simd}{', avx512_bf16' if 'avx512_bf16' in flags else ''}, "
            f"{os.cpu_count()} logical CPUs visible)")


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


def time_matvec(fn, v, u, steps, warmup, torch):
    s = torch.cuda.current_stream()
    for _ in range(warmup):
        fn(v, u, s)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(s)
    for _ in range(steps):
        fn(v, u, s)
    e1.record(s)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def host_boundary(op, u_dev, reps: int, torch) -> dict:
    """MatMult_mf on host arrays (include/aMatFree.hpp:3893-3941, the PETSc shell
    hook's form): copy u in, the matvec, copy v out, synchronized; pageable numpy
    arrays and pinned ones.  Reported beside the device rate, never as `value`."""
    import numpy as np
    n = u_dev.numel()
    out = {}
    for kind in ("pageable", "pinned"):
        if kind == "pinned":
            tu = torch.empty(n, dtype=torch.float64, pin_memory=True)
            tv = torch.empty(n, dtype=torch.float64, pin_memory=True)
            u, v = tu.numpy(), tv.numpy()
        else:
            u, v = np.empty(n), np.empty(n)
        u[:] = u_dev.cpu().numpy()
        for _ in range(2):
            op.MatMult_mf(u, out=v)
        t0 = time.perf_counter()
        for _ in range(reps):
            op.MatMult_mf(u, out=v)
        dt = (time.perf_counter() - t0) / reps
        out[kind] = {"ms_per_matvec": dt * 1e3, "GDoF_s": n / dt / 1e9}
    return out


def precision_mode(args, lite: bool = False) -> dict:
    """Precision: jittered tet10 Poisson (n^3 hexes x 6 tets), b = M f, identity
    BCs; FP64 CG vs FP64 iterative refinement with FP32 / BF16 element
    matrices (restarted per outer step, and with reliable updates), all to
    1e-10 relative residual, error against the FP64 solution.  lite: the
    reliable-update forms only, no method-2 operator."""
    import numpy as np
    import torch
    from amat_amd import BCOP, PREC, Maps, aMatFree, problems
    t0 = time.time()
    p = problems.poisson_tet10(args.n, seed=3, jitter=0.2)
    log(f"[precision] tet10 {args.n}^3 x 6: {p.n_elem} elements, {p.n_dofs} DoFs ({time.time() - t0:.1f}s)")
    op = p.build_operator(aMatFree, Maps, precision=PREC.FP64 | PREC.FP32 | PREC.BF16, bcop=BCOP.IDENTITY,
                          sym_storage=args.sym_storage)
    del p.K
    log(f"[precision] operator finalized ({time.time() - t0:.1f}s)")
    b = op.apply_bc(p.rhs)
    db = torch.from_numpy(b).cuda()
    s = torch.cuda.current_stream()
    inf = op.info()
    res = {"metric": "precision: IR-CG vs FP64 CG to 1e-10, jittered tet10 Poisson", "elements": p.n_elem, "dofs": p.n_dofs,
           "packed_symmetric": bool(inf.packed_symmetric)}
    # matvec rates per storage; algorithmic bytes of the stored form (n^2 values
    # per element, or n(n+1)/2 for packed symmetric tiles)
    E, n, N = p.n_elem, 10, p.n_dofs
    nv = n * (n + 1) // 2 if inf.packed_symmetric else n * n
    u64 = torch.from_numpy(p.random_u()).cuda()
    v64 = torch.empty_like(u64)
    u32 = u64.float()
    v32 = torch.empty_like(u32)
    for name, fn, v, u, sk, sv in (("fp64", op.matmult_dev, v64, u64, 8, 8),
                                   ("fp32", lambda a, b_, st: op.matmult_f32_dev(PREC.FP32, a, b_, st), v32, u32, 4, 4),
                                   ("bf16", lambda a, b_, st: op.matmult_f32_dev(PREC.BF16, a, b_, st), v32, u32, 2, 4)):
        ms = time_matvec(fn, v, u, args.steps, args.warmup, torch)
        diag_bytes = E * n * 4 if name == "bf16" else 0
        alg = E * (nv * sk + 4 * n) + diag_bytes + N * 2 * sv
        full = E * (n * n * sk + 4 * n) + diag_bytes + N * 2 * sv
        res[f"matvec_{name}"] = {"ms": ms, "GDoF/s": N / ms / 1e6, "hbm_frac": alg / (ms * 1e-3) / HBM_PEAK,
                                 "alg_bytes": alg, "whole_matrix_bytes": full}
        log(f"[precision] matvec {name}: {ms:.3f} ms, {N / ms / 1e6:.2f} GDoF/s")
    x64 = torch.zeros_like(db)
    torch.cuda.synchronize()
    t = time.perf_counter()
    st = op.cg(db, x64, rtol=1e-10, max_it=20000, stream=s)
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t
    res["cg_fp64"] = {"seconds": t64, "iterations": st.iterations, "rel_residual": st.rel_residual,
                      "converged": bool(st.converged)}
    log(f"[precision] FP64 CG: {st.iterations} it, {t64:.3f} s, relres {st.rel_residual:.2e}")
    xr = x64.cpu().numpy()
    from amat_amd.amat import IR_RELIABLE
    # restarted refinement (inner tolerance --inner-rtol) and reliable updates
    # (replacement threshold --ru-delta)
    forms = ((PREC.FP32, True), (PREC.BF16, True)) if lite else \
        ((PREC.FP32, False), (PREC.BF16, False), (PREC.FP32, True), (PREC.BF16, True))
    for prec, reliable in forms:
        x = torch.zeros_like(db)
        torch.cuda.synchronize()
        t = time.perf_counter()
        si = op.ir_cg(int(prec) | (IR_RELIABLE if reliable else 0), db, x, rtol=1e-10,
                      inner_rtol=args.ru_delta if reliable else args.inner_rtol, max_outer=200 if reliable else 60,
                      max_inner=20000, stream=s)
        torch.cuda.synchronize()
        ti = time.perf_counter() - t
        err = float(np.linalg.norm(x.cpu().numpy() - xr) / np.linalg.norm(xr))
        key = f"{'ru' if reliable else 'ir'}_cg_{prec.name.lower()}"
        res[key] = {"seconds": ti, "outer": si.outer_iterations, "inner": si.iterations,
                    "rel_residual": si.rel_residual, "converged": bool(si.converged),
                    "rel_error_vs_fp64_solution": err, "speedup_vs_fp64_cg": t64 / ti}
        log(f"[precision] {'RU' if reliable else 'IR'}-CG {prec.name}: {'replacements' if reliable else 'outer'} "
            f"{si.outer_iterations} inner {si.iterations}, {ti:.3f} s, relres {si.rel_residual:.2e}, "
            f"err vs FP64 {err:.2e}, speedup {t64 / ti:.2f}x")
    if lite:
        return res
    # method 2 (FREE): the same operator recomputed in-kernel from the coordinates,
    # FP32 inner solves with FP64 reliable updates (nothing stored per precision)
    del op
    torch.cuda.empty_cache()
    fr = p.build_free_operator(aMatFree, Maps, bcop=BCOP.IDENTITY)
    ms = time_matvec(lambda a, b_, st_: fr.matmult_f32_dev(PREC.FP32, a, b_, st_), v32, u32, args.steps, args.warmup,
                     torch)
    x = torch.zeros_like(db)
    torch.cuda.synchronize()
    t = time.perf_counter()
    si = fr.ir_cg(int(PREC.FP32) | IR_RELIABLE, db, x, rtol=1e-10, inner_rtol=args.ru_delta, max_outer=200,
                  max_inner=20000, stream=s)
    torch.cuda.synchronize()
    ti = time.perf_counter() - t
    err = float(np.linalg.norm(x.cpu().numpy() - xr) / np.linalg.norm(xr))
    res["ru_cg_fp32_free"] = {"matvec_ms": ms, "seconds": ti, "outer": si.outer_iterations, "inner": si.iterations,
                              "rel_residual": si.rel_residual, "converged": bool(si.converged),
                              "rel_error_vs_fp64_solution": err, "speedup_vs_fp64_cg": t64 / ti}
    log(f"[precision] method-2 FP32 matvec {ms:.3f} ms; RU-CG FP32 on method 2: replacements {si.outer_iterations} "
        f"inner {si.iterations}, {ti:.3f} s, relres {si.rel_residual:.2e}, err vs FP64 {err:.2e}, "
        f"speedup {t64 / ti:.2f}x")
    return res


FP64_PEAK = 78.6e12  # MI355X FP64 (vector and matrix), flop/s


def free_mode(args) -> dict:
    """Method 2 (FREE) on the headline workload: hex20 elasticity n^3, the same
    operator with K_e computed in-kernel from the node coordinates at every
    application (no element matrices in HBM), against the stored-matrix FP64
    operator; Gauss rules 2^3 (the reference example's NGT) and 3^3."""
    import torch
    from amat_amd import BCOP, PREC, Maps, aMatFree, problems
    res = {"metric": "method 2 (FREE) matvec vs stored K_e (FP64; FP32 forms for hex20)", "n": args.n}
    s = torch.cuda.current_stream()
    for ngt in (3, 2):
        t0 = time.time()
        p = problems.elasticity_hex20(args.n, ngt=ngt)
        st = p.build_operator(aMatFree, Maps, precision=PREC.FP64 | PREC.FP32, bcop=BCOP.PENALTY_DIAG, penalty=1e8)
        fr = p.build_free_operator(aMatFree, Maps, bcop=BCOP.PENALTY_DIAG, penalty=1e8)
        log(f"[free] hex20 {args.n}^3 ngt {ngt}: {p.n_elem} elements, {p.n_dofs} DoFs, operators built "
            f"({time.time() - t0:.1f}s)")
        u = torch.from_numpy(p.random_u()).cuda()
        v0, v1 = torch.empty_like(u), torch.empty_like(u)
        st.matvec_dev(v0, u, s)   # K u without the BC rows, which the penalty would dominate
        fr.matvec_dev(v1, u, s)
        torch.cuda.synchronize()
        rel = float(torch.linalg.norm(v1 - v0) / torch.linalg.norm(v0))
        N, E = p.n_dofs, p.n_elem
        for name, op in (("stored", st), ("free", fr)):
            ms = time_matvec(op.matmult_dev, v0, u, args.steps, args.warmup, torch)
            ms_e, ms_g, _ = op.time_kernels_dev(args.steps, v0, u, s)
            inf = op.info()
            r = {"ms": ms, "GDoF/s": N / ms / 1e6, "element_kernel_ms": ms_e, "gather_ms": ms_g,
                 "device_bytes": inf.device_bytes}
            if name == "free":
                r["flops_per_matvec"] = inf.flops
                r["element_kernel_TFLOP/s"] = inf.flops / (ms_e * 1e-3) / 1e12
                r["fp64_peak_frac"] = inf.flops / (ms_e * 1e-3) / FP64_PEAK
                r["rel_l2_vs_stored"] = rel
            else:
                elem_bytes = E * (60 * 60 * 8 + 4 * 60) + N * 8
                r["element_kernel_hbm_frac"] = elem_bytes / (ms_e * 1e-3) / HBM_PEAK
            res[f"ngt{ngt}_{name}"] = r
            log(f"[free] ngt {ngt} {name}: {ms:.4f} ms/matvec ({N / ms / 1e6:.2f} GDoF/s), element kernel "
                f"{ms_e:.4f} ms, gather {ms_g:.4f} ms, device {inf.device_bytes / 1e9:.2f} GB"
                + (f", {inf.flops / (ms_e * 1e-3) / 1e12:.2f} TFLOP/s, rel L2 vs stored {rel:.2e}"
                   if name == "free" else ""))
        res[f"ngt{ngt}_speedup"] = res[f"ngt{ngt}_stored"]["ms"] / res[f"ngt{ngt}_free"]["ms"]
        # FP32 forms (the inner operator of the mixed-precision solvers): FP32
        # element matrices vs the FREE kernels in FP32 arithmetic
        u32 = u.float()
        v32 = torch.empty_like(u32)
        fr.matmult_f32_dev(PREC.FP32, v32, u32, s)
        fr.matmult_dev(v1, u, s)
        torch.cuda.synchronize()
        rel32 = float(torch.linalg.norm(v32.double() - v1) / torch.linalg.norm(v1))
        for name, op in (("stored_fp32", st), ("free_fp32", fr)):
            ms = time_matvec(lambda a, b_, st_, o=op: o.matmult_f32_dev(PREC.FP32, a, b_, st_), v32, u32, args.steps,
                             args.warmup, torch)
            res[f"ngt{ngt}_{name}"] = {"ms": ms, "GDoF/s": N / ms / 1e6}
            if name == "free_fp32":
                res[f"ngt{ngt}_{name}"]["rel_l2_vs_free_fp64"] = rel32
            log(f"[free] ngt {ngt} {name}: {ms:.4f} ms/matvec ({N / ms / 1e6:.2f} GDoF/s)"
                + (f", rel L2 vs the FP64 form {rel32:.2e}" if name == "free_fp32" else ""))
        del st, fr, u, v0, v1, u32, v32
        torch.cuda.empty_cache()
    # Laplace kinds, one lane per element: hex8 Poisson (2x2x2 Gauss) at (2n)^3,
    # jittered tet10 Poisson (8-point rule) at n^3 x 6
    for label, p in ((f"hex8_{2 * args.n}", problems.poisson_hex8(2 * args.n)),
                     (f"tet10_{args.n}", problems.poisson_tet10(args.n, with_rhs=False))):
        st = p.build_operator(aMatFree, Maps, bcop=BCOP.IDENTITY)
        fr = p.build_free_operator(aMatFree, Maps, bcop=BCOP.IDENTITY)
        u = torch.from_numpy(p.random_u()).cuda()
        v0, v1 = torch.empty_like(u), torch.empty_like(u)
        st.matvec_dev(v0, u, s)
        fr.matvec_dev(v1, u, s)
        torch.cuda.synchronize()
        rel = float(torch.linalg.norm(v1 - v0) / torch.linalg.norm(v0))
        for name, op in (("stored", st), ("free", fr)):
            ms = time_matvec(op.matmult_dev, v0, u, args.steps, args.warmup, torch)
            ms_e, ms_g, _ = op.time_kernels_dev(args.steps, v0, u, s)
            inf = op.info()
            r = {"ms": ms, "GDoF/s": p.n_dofs / ms / 1e6, "element_kernel_ms": ms_e, "gather_ms": ms_g,
                 "device_bytes": inf.device_bytes, "alg_bytes_per_matvec": inf.alg_bytes_fp64,
                 "hbm_frac": inf.alg_bytes_fp64 / (ms * 1e-3) / HBM_PEAK}
            if name == "free":
                r["rel_l2_vs_stored"] = rel
                r["element_kernel_TFLOP/s"] = inf.flops / (ms_e * 1e-3) / 1e12
            res[f"{label}_{name}"] = r
            log(f"[free] {label} {name}: {ms:.4f} ms/matvec ({p.n_dofs / ms / 1e6:.2f} GDoF/s), element kernel "
                f"{ms_e:.4f} ms, gather {ms_g:.4f} ms" +
                (f", {r['element_kernel_TFLOP/s']:.2f} TFLOP/s, rel L2 vs stored {rel:.2e}" if name == "free" else ""))
        del st, fr, u, v0, v1, p
        torch.cuda.empty_cache()
    return res


MFMA_PEAK = {"fp32": 157.3e12, "bf16": 2.5e15}  # MI355X dense MFMA, flop/s


def batch_mode(args) -> dict:
    """Batching: multi-RHS products of hex20 elasticity 48^3 and hex8 Poisson
    80x40x20 together in one persistent MFMA launch, nrhs in {8, 32, 128},
    FP32 and BF16 storage, FP32 vectors (n_total x nrhs, DoF-major).  Per case:
    autotuned tile shape, mean ms per launch (eager and HIP-graph replay),
    algorithmic bytes and flops, and the error of RHS column 0 against the FP64
    operator."""
    import numpy as np
    import torch
    from amat_amd import PREC, SCATTER, Maps, MultiRHS, aMatFree, mesh, problems
    t0 = time.time()
    probs = [problems.elasticity_hex20(args.batch_n), problems.poisson_hex8(80, 40, 20)]
    ops = [p.build_operator(aMatFree, Maps, precision=PREC.FP64) for p in probs]
    log(f"[batch] operators: " + ", ".join(f"{p.name} E={p.n_elem} N={p.n_dofs}" for p in probs) +
        f" ({time.time() - t0:.1f}s)")
    s = torch.cuda.Stream()
    res = {"metric": "multi-RHS EbE products (hex20 48^3 + hex8 80x40x20, one persistent launch)",
           "operators": [{"name": p.name, "elements": p.n_elem, "dofs": p.n_dofs, "dofs_per_element": p.dofs_per_elem}
                         for p in probs], "cases": []}
    scatters = {"atomic": SCATTER.ATOMIC, "deterministic": SCATTER.DETERMINISTIC}
    for prec_name, sc_name, nrhs in [(p_, s_, int(n_)) for p_ in args.batch_precisions.split(",")
                                     for s_ in args.batch_scatter.split(",") for n_ in args.batch_nrhs.split(",")]:
        prec = PREC.FP32 if prec_name == "fp32" else PREC.BF16
        mr = MultiRHS(ops, prec)
        mr.set_scatter(scatters[sc_name])
        sk = 4 if prec == PREC.FP32 else 2
        U = [torch.from_numpy(mesh.uniform(21 + i, 7, p.n_dofs * nrhs, -1.0, 1.0).astype(np.float32)
                              .reshape(p.n_dofs, nrhs)).cuda() for i, p in enumerate(probs)]
        V = [torch.empty_like(u) for u in U]
        with torch.cuda.stream(s):
            if args.batch_tiles:   # a fixed launch shape (profiling passes): no sweep
                mr.set_tiles(*[int(x) for x in args.batch_tiles.split(",")])
                tunes = {}
            else:
                tunes = mr.autotune(nrhs, U, V, reps=5, stream=s)
            for _ in range(args.warmup):
                mr.apply_dev(nrhs, U, V, s)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            s.synchronize()
            e0.record(s)
            for _ in range(args.steps):
                mr.apply_dev(nrhs, U, V, s)
            e1.record(s)
            s.synchronize()
            ms = e0.elapsed_time(e1) / args.steps
            ms_graph = float("nan")
            if not args.batch_tiles:
                mr.graph_capture(nrhs, U, V, s)
                mr.graph_launch(s)
                s.synchronize()
                e0.record(s)
                for _ in range(args.steps):
                    mr.graph_launch(s)
                e1.record(s)
                s.synchronize()
                ms_graph = e0.elapsed_time(e1) / args.steps
        # column 0 against the FP64 operator (plain element sum)
        errs = []
        for i, (p, op) in enumerate(zip(probs, ops)):
            u0 = U[i][:, 0].double().contiguous()
            v0 = torch.empty_like(u0)
            op.matvec_dev(v0, u0, torch.cuda.current_stream())
            errs.append(float(torch.linalg.norm(V[i][:, 0].double() - v0) / torch.linalg.norm(v0)))
        alg = sum(p.n_elem * (p.dofs_per_elem ** 2 * sk + 4 * p.dofs_per_elem) + p.n_dofs * 8 * nrhs
                  for p in probs)
        flops = sum(2 * p.n_elem * p.dofs_per_elem ** 2 * nrhs for p in probs)
        best = min(ms, ms_graph) if ms_graph == ms_graph else ms
        case = {"precision": prec_name, "scatter": sc_name, "nrhs": nrhs, "tile": list(mr.tiles()), "ms": ms,
                "ms_graph": ms_graph,
                "GDoF_rhs_per_s": sum(p.n_dofs for p in probs) * nrhs / (best * 1e-3) / 1e9,
                "alg_bytes": alg, "hbm_frac": alg / (best * 1e-3) / HBM_PEAK,
                "tflops": flops / (best * 1e-3) / 1e12, "mfma_frac": flops / (best * 1e-3) / MFMA_PEAK[prec_name],
                "col0_rel_err_vs_fp64": errs,
                "autotune_ms": {f"{rt}x{w}x{c}": round(t, 4) for (rt, w, c), t in tunes.items()}}
        res["cases"].append(case)
        log(f"[batch] {prec_name} {sc_name} nrhs {nrhs}: tile {mr.tiles()} {ms:.4f} ms (graph {ms_graph:.4f}), "
            f"{case['hbm_frac'] * 100:.1f}% HBM, {case['tflops']:.1f} TFLOP/s, err {errs}")
        del U, V
        del mr
    return res


def extras(args) -> dict:
    """The precision, batching and serving workloads measured in the default
    run, in brief: jittered tet10 Poisson 96^3 x 6 (matvec per storage, FP64 CG
    against FP32 / BF16 storage with FP64 reliable updates, to 1e-10), the
    multi-RHS products of hex20 48^3 + hex8 80x40x20 (best autotuned launch
    per storage and nrhs over both scatter forms), and a served stream of 400
    CG solve requests at 800 req/s (FP64 CG and FP32 storage with reliable
    updates on stored operators, FP32 inner solves on method-2 operators)."""
    import copy
    import torch
    out = {}
    a = copy.copy(args)
    a.n, a.steps, a.warmup = 96, min(args.steps, 20), 3
    t0 = time.time()
    pr = precision_mode(a, lite=True)
    torch.cuda.empty_cache()
    keep = ("seconds", "iterations", "inner", "outer", "rel_residual", "converged", "rel_error_vs_fp64_solution",
            "speedup_vs_fp64_cg")
    # hbm_frac: algorithmic bytes of the stored form (packed triangles when
    # packed_symmetric) over the whole matvec; whole_matrix_frac: the same time
    # against the bytes of whole n x n element matrices
    out["precision_tet10_96"] = {
        "elements": pr["elements"], "dofs": pr["dofs"], "packed_symmetric": pr["packed_symmetric"],
        "matvec": {k[7:]: {"ms": v["ms"], "hbm_frac": v["hbm_frac"], "alg_bytes": v["alg_bytes"],
                           "whole_matrix_frac": v["whole_matrix_bytes"] / (v["ms"] * 1e-3) / HBM_PEAK}
                   for k, v in pr.items() if k.startswith("matvec_")},
        **{k: {f: v[f] for f in keep if f in v} for k, v in pr.items() if k.startswith(("cg_", "ru_cg"))},
        "seconds_in_bench": time.time() - t0}
    t0 = time.time()
    b = copy.copy(args)
    b.steps, b.warmup = min(args.steps, 20), 3
    br = batch_mode(b)
    torch.cuda.empty_cache()
    best = {}
    for c in br["cases"]:
        key = f"{c['precision']}_nrhs{c['nrhs']}"
        ms = min(c["ms"], c["ms_graph"])
        if key not in best or ms < best[key]["ms"]:
            best[key] = {"ms": ms, "scatter": c["scatter"], "tile": c["tile"], "hbm_frac": c["hbm_frac"],
                         "tflops": c["tflops"], "col0_rel_err_vs_fp64": c["col0_rel_err_vs_fp64"]}
    out["batch_hex20_48_hex8_80x40x20"] = {"best": best, "seconds_in_bench": time.time() - t0}
    t0 = time.time()
    c = copy.copy(args)
    c.serve_requests, c.serve_rates, c.serve_methods = 400, "800", "fp64,fp32r,fp32rfree"
    sr = serve_mode(c)
    torch.cuda.empty_cache()
    keep = ("solves_per_s", "p50_s", "p99_s", "service_p50_s", "service_p99_s", "iterations", "all_converged",
            "max_rel_residual", "speedup_same_form", "speedup_vs_stored_fp64")
    out["serve_400_at_800"] = {"requests": sr["requests"], "operators": sr["operators"],
                               **{r["method"]: {k: r[k] for k in keep if k in r} for r in sr["runs"]},
                               "seconds_in_bench": time.time() - t0}
    return out


def serve_mode(args) -> dict:
    """Serving: a stream of CG solve requests (60% hex8 Poisson n~U{16..96},
    30% hex20 elasticity n~U{8..40}, 10% tet10 Poisson n~U{12..48}, own RHS,
    1e-10 relative residual) arriving as Poisson traffic, served by the
    continuous-batching server under an HBM budget; FP64 CG and FP64-refined
    FP32/BF16 CG on the same request streams.  Operators are built once
    (resident cache) before the timed streams, or, with --serve-cold, inside
    them on first use (and again after an eviction when the budget binds)."""
    import numpy as np
    from amat_amd import PREC, serve
    from amat_amd.amat import IR_RELIABLE
    rates = [float(r) for r in args.serve_rates.split(",")]
    # fp32 / bf16: refinement restarted per outer step (inner tolerance --inner-rtol);
    # fp32r / bf16r: reliable updates (replacement threshold --ru-delta)
    # fp64free: FP64 CG on method-2 operators (K_e recomputed in-kernel; own cache);
    # fp32rfree: the same operators, FP32 inner solves with FP64 reliable updates
    methods = {"fp64": PREC.FP64, "fp32": PREC.FP32, "bf16": PREC.BF16,
               "fp32r": int(PREC.FP32) | IR_RELIABLE, "bf16r": int(PREC.BF16) | IR_RELIABLE, "fp64free": PREC.FP64,
               "fp32rfree": int(PREC.FP32) | IR_RELIABLE}
    reliable = {"fp32r", "bf16r", "fp32rfree"}
    names = args.serve_methods.split(",")
    streams = {r: serve.make_requests(args.serve_requests, r, seed=args.serve_seed) for r in rates}
    keys = sorted({(q.kind, q.n) for st in streams.values() for q in st})
    budget = int(args.serve_budget_gb * 2**30)
    caches = {}
    for free in sorted({n.endswith("free") for n in names}):
        cache = serve.OperatorCache(free=free)
        t0 = time.time()
        for i, key in enumerate(keys if not args.serve_cold else []):
            cache.build(key)
            if i % 10 == 0:
                log(f"[serve] built {i + 1}/{len(keys)} {'method-2 ' if free else ''}operators "
                    f"({cache.bytes / 2**30:.1f} GiB, {time.time() - t0:.0f}s)")
        log(f"[serve] {len(keys)} {'method-2 ' if free else ''}operators resident: {cache.bytes / 2**30:.1f} GiB, "
            f"built in {cache.build_seconds:.1f}s")
        caches[free] = cache
    cache = caches.get(False, caches.get(True))
    res = {"metric": "served CG solves/s and latency at 1e-10 relative residual", "requests": args.serve_requests,
           "rates": rates, "budget_GiB": args.serve_budget_gb, "slots": args.serve_slots, "step_iters": args.serve_iters,
           "inner_rtol": args.inner_rtol, "ru_delta": args.ru_delta, "cold_start": bool(args.serve_cold),
           "admission": args.serve_admission, "build_workers": args.serve_build_workers,
           "operators": len(keys), "operator_GiB": cache.bytes / 2**30, "operator_build_s": cache.build_seconds,
           "runs": []}
    for rate in rates:
        for name in names:
            server = serve.Server(max_jobs=args.serve_slots)
            loop = serve.ServingLoop(server, caches[name.endswith("free")], methods[name], budget, n_iter=args.serve_iters,
                                     inner_rtol=args.ru_delta if name in reliable else args.inner_rtol,
                                     admission=args.serve_admission, build_workers=args.serve_build_workers)
            last = [0.0]

            def prog(el, fin, run, pend):
                if el - last[0] >= 10.0:
                    last[0] = el
                    log(f"[serve] rate {rate:g} {name}: {el:.0f}s finished {fin} running {run} waiting {pend}")

            m = loop.run(streams[rate], progress=prog)
            done = m.pop("done")
            by_kind = {}
            for k in ("hex8", "hex20", "tet10"):
                lat = [d.finished - d.arrival for d in done if d.kind == k]
                if lat:
                    by_kind[k] = {"count": len(lat), "p50_s": float(np.percentile(lat, 50)),
                                  "p99_s": float(np.percentile(lat, 99))}
            m.update({"rate": rate, "method": name, "by_kind": by_kind,
                      "iterations": int(sum(d.iterations for d in done))})
            res["runs"].append(m)
            log(f"[serve] rate {rate:g} {name}: {m['solves_per_s']:.1f} solves/s, latency p50 {m['p50_s'] * 1e3:.0f} ms "
                f"p99 {m['p99_s'] * 1e3:.0f} ms, service p50 {m['service_p50_s'] * 1e3:.0f} ms p99 "
                f"{m['service_p99_s'] * 1e3:.0f} ms, builds {m['operator_builds']} ({m['operator_build_s']:.1f}s), "
                f"evictions {m['evictions']}, budget waits {m['budget_waits']}, converged {m['all_converged']}, "
                f"max relres {m['max_rel_residual']:.2e}")
            del server
    # same-form speedups (stored K: against FP64 CG on stored K; method 2: against FP64
    # CG on the method-2 operators) and the cross-form one against stored FP64 CG
    for rate in rates:
        runs = {r["method"]: r for r in res["runs"] if r["rate"] == rate}
        for name, r in runs.items():
            same = runs.get("fp64free" if name.endswith("free") else "fp64")
            if same:
                r["speedup_same_form"] = r["solves_per_s"] / same["solves_per_s"]
            if "fp64" in runs:
                r["speedup_vs_stored_fp64"] = r["solves_per_s"] / runs["fp64"]["solves_per_s"]
    return res


def calibrate_copy(torch) -> None:
    """1 GiB copied device to device 5 times: known bytes for the FETCH_SIZE /
    WRITE_SIZE counter scale of a profiling pass."""
    src = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--mode", choices=["matvec", "precision", "batch", "serve", "free"], default="matvec")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-rates", default="50,200,800")
    ap.add_argument("--serve-methods", default="fp64,fp32r,fp32,bf16")
    ap.add_argument("--ru-delta", type=float, default=0.1,
                    help="reliable-update replacement threshold (methods fp32r, bf16r, --mode precision)")
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--serve-cold", action="store_true",
                    help="start with no resident operator: builds, admission waits and evictions happen in the stream")
    ap.add_argument("--serve-slots", type=int, default=64)
    ap.add_argument("--serve-admission", choices=["fifo", "affinity"], default="fifo",
                    help="admission order: oldest first, or oldest with a resident operator first (bounded bypass)")
    ap.add_argument("--serve-build-workers", type=int, default=0,
                    help="threads building operators while the server steps (0: built inline)")
    ap.add_argument("--serve-iters", type=int, default=8)
    ap.add_argument("--serve-seed", type=int, default=7)
    ap.add_argument("--batch-n", type=int, default=48, help="hex20 elements per edge in --mode batch")
    ap.add_argument("--batch-nrhs", default="8,32,128")
    ap.add_argument("--batch-precisions", default="fp32,bf16")
    ap.add_argument("--batch-scatter", default="atomic,deterministic")
    ap.add_argument("--batch-tiles", default="",
                    help="rhs_tile,waves,split: a fixed launch shape instead of the autotuner (profiling passes)")
    ap.add_argument("--inner-rtol", type=float, default=1e-2)
    ap.add_argument("--sym-storage", type=int, choices=[0, 1], default=None,
                    help="element-matrix storage of 8-/10-DoF groups in --mode precision: 0 whole, 1 packed when "
                         "symmetric (default: the library's, 1)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=64, help="elements per cube edge")
    ap.add_argument("--scatter", choices=["deterministic", "atomic"], default="deterministic")
    ap.add_argument("--cpu-layers", type=int, default=1 << 30, help="element layers in the CPU baseline sample (all)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="CPU time of each baseline variant")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the precision (tet10 96^3) and batching (multi-RHS) workloads of the default run")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after the timed region, copy 1 GiB device-to-device 5 times (known bytes for the "
                         "FETCH_SIZE / WRITE_SIZE counter scale)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU (one rank's share of a partitioned mesh); --gpus {args.gpus} is not "
              "supported: multi-GPU runs are out of scope", file=sys.stderr)
        return 2

    import torch
    from amat_amd import BCOP, PREC, SCATTER, Maps, aMatFree, problems

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    if args.mode == "precision":
        print(json.dumps(precision_mode(args)), flush=True)
        if args.calibrate_copy:
            calibrate_copy(torch)
        return 0
    if args.mode == "batch":
        print(json.dumps(batch_mode(args)), flush=True)
        if args.calibrate_copy:
            calibrate_copy(torch)
        return 0
    if args.mode == "serve":
        print(json.dumps(serve_mode(args)), flush=True)
        return 0
    if args.mode == "free":
        print(json.dumps(free_mode(args)), flush=True)
        return 0
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
        calibrate_copy(torch)
    # measured peak beside the nominal one: a 16-byte streaming read of 2 GiB
    from amat_amd.amat import measure_read_bandwidth
    read_bw = measure_read_bandwidth(torch.cuda.current_device(), 2 << 30, 10)
    log(f"[bench] measured HBM read stream: {read_bw / 1e9:.0f} GB/s")
    host = host_boundary(op, u, args.steps, torch)
    log(f"[bench] host-array MatMult_mf (PCIe copies included): pageable {host['pageable']['ms_per_matvec']:.3f} ms, "
        f"pinned {host['pinned']['ms_per_matvec']:.3f} ms")
    E, n, N = p.n_elem, p.dofs_per_elem, p.n_dofs
    # K as stored (near-symmetric records or whole matrices) + map + u read
    elem_bytes = info.k_bytes_fp64 + E * 4 * n + N * 8
    whole_bytes = E * (n * n * 8 + 4 * n) + N * 8       # the same with whole n x n FP64 matrices
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
                     "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "kernel": "ebe_nsym_kernel<60>" if info.near_symmetric else "ebe_tile_kernel<60,double>",
                     "near_symmetric_records": bool(info.near_symmetric),
                     "measured_read_GBps": read_bw / 1e9, "frac_of_measured_read": achieved / read_bw,
                     "alg_bytes_per_launch": elem_bytes, "whole_matrix_bytes_per_launch": whole_bytes,
                     "whole_matrix_equivalent_GBps": whole_bytes / (ms_elem * 1e-3) / 1e9,
                     "ms_per_launch": ms_elem, "gather_ms": ms_gather},
        "host_arrays_pcie_inclusive": host,
    }
    if not args.no_extra:
        del op, u, v
        torch.cuda.empty_cache()
        out["extra"] = extras(args)
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log(f"[bench] CPU baseline on {threads} threads ...")
        out["cpu_baseline"] = cpu_baseline(p, args.cpu_layers, args.cpu_seconds, threads)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
