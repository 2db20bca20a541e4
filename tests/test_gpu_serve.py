"""GPU tests of the continuous-batching CG server (amx_server_*).

Bar: a job's solution is BIT-IDENTICAL alone, in a mix admitted together, in
a mix where it joins while others are iterating, and to the single-solve entry
points (amx_cg_solve_dev for FP64 CG, amx_ir_cg_solve_dev for refinement):
the batched launches run each job's own fixed-order arithmetic (the
reliable-update form stops its recurrence on the device, so the host's poll
schedule -- 8 iterations per step alone, 4 in the join test, the
single-solve's own chunk -- does not change the iterates).  Every served
solve reaches 1e-10 true relative residual (FP64, measured on the device),
and the FP64 CG solution agrees with scipy's sparse direct solve of the
oracle-assembled system to 1e-8.
"""
import numpy as np
import pytest

from amat_amd import PREC, AmatError, serve
from amat_amd.amat import IR_RELIABLE
from amat_amd.serve import Server, build_operator, build_problem, rhs_for

pytestmark = pytest.mark.gpu

CASES = [("hex8", 7), ("hex20", 3), ("tet10", 3), ("hex8", 5)]
# FP64 CG; refinement restarted per outer step; refinement with reliable
# updates (FP32: BF16's operator error stalls the unrestarted recurrence on
# elasticity, so BF16 reliable updates are tested on the Laplace cases)
FP32R = int(PREC.FP32) | IR_RELIABLE
BF16R = int(PREC.BF16) | IR_RELIABLE
METHODS = [PREC.FP64, PREC.FP32, PREC.BF16, FP32R]


@pytest.fixture(scope="module")
def ops(torch_gpu):
    out = []
    for i, (kind, n) in enumerate(CASES):
        p = build_problem(kind, n)
        out.append((p, build_operator(p), rhs_for(p, 100 + i)))
    return out


def solve_alone(op, p, b, method):
    s = Server(max_jobs=4)
    jid = s.submit(op, method, b)
    while True:
        s.step(8)
        ids = s.poll()
        if ids:
            assert ids == [jid]
            return s.result(jid, p.n_dofs)


def run_all(server, jobs, n):
    """step until every job id in `jobs` is reported; returns {id: (x, stats)}."""
    out = {}
    while len(out) < len(jobs):
        server.step(8)
        for jid in server.poll():
            out[jid] = server.result(jid, n[jid])
    return out


@pytest.mark.parametrize("method", METHODS)
def test_alone_equals_mix(ops, method):
    alone = [solve_alone(op, p, b, method) for p, op, b in ops]
    for x, st in alone:
        assert st.converged and st.rel_residual <= 1e-10
    s = Server(max_jobs=8)
    ids = [s.submit(op, method, b) for p, op, b in ops]
    res = run_all(s, ids, {j: p.n_dofs for j, (p, _, _) in zip(ids, ops)})
    for jid, (xa, sa) in zip(ids, alone):
        xm, sm = res[jid]
        np.testing.assert_array_equal(xm, xa)
        assert sm.iterations == sa.iterations and sm.outer_iterations == sa.outer_iterations


@pytest.mark.parametrize("method", METHODS)
def test_join_while_running(ops, method):
    """Continuous batching: jobs admitted at different steps, while others
    iterate, still give the alone results."""
    alone = [solve_alone(op, p, b, method) for p, op, b in ops]
    s = Server(max_jobs=8)
    ids, sizes, out = [], {}, {}
    for k, (p, op, b) in enumerate(ops):
        jid = s.submit(op, method, b)
        ids.append(jid)
        sizes[jid] = p.n_dofs
        for _ in range(k + 1):  # a few steps before the next arrival
            s.step(4)
            for j in s.poll():
                out[j] = s.result(j, sizes[j])
    while len(out) < len(ids):
        s.step(4)
        for j in s.poll():
            out[j] = s.result(j, sizes[j])
    for jid, (xa, _) in zip(ids, alone):
        np.testing.assert_array_equal(out[jid][0], xa)


def test_matches_single_solve_entry_points(ops, torch_gpu):
    torch = torch_gpu
    for p, op, b in ops:
        db = torch.from_numpy(b).cuda()
        x = torch.zeros_like(db)
        st = op.cg(db, x, rtol=1e-10, max_it=20000)
        xs, ss = solve_alone(op, p, b, PREC.FP64)
        np.testing.assert_array_equal(xs, x.cpu().numpy())
        assert ss.iterations == st.iterations
        for prec in (PREC.FP32, PREC.BF16, FP32R) + ((BF16R,) if p.name != "elasticity_hex20" else ()):
            x = torch.zeros_like(db)
            st = op.ir_cg(prec, db, x, rtol=1e-10, inner_rtol=1e-2, max_outer=60, max_inner=20000)
            xs, ss = solve_alone(op, p, b, prec)
            np.testing.assert_array_equal(xs, x.cpu().numpy())
            assert ss.iterations == st.iterations and ss.outer_iterations == st.outer_iterations


def test_bf16_reliable_laplace_mix(ops):
    lap = [(p, op, b) for p, op, b in ops if p.name != "elasticity_hex20"]
    alone = [solve_alone(op, p, b, BF16R) for p, op, b in lap]
    for x, st in alone:
        assert st.converged and st.rel_residual <= 1e-10
    s = Server(max_jobs=8)
    ids = [s.submit(op, BF16R, b) for p, op, b in lap]
    res = run_all(s, ids, {j: p.n_dofs for j, (p, _, _) in zip(ids, lap)})
    for jid, (xa, sa) in zip(ids, alone):
        np.testing.assert_array_equal(res[jid][0], xa)


def test_fp64_solution_vs_direct(ops):
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    for p, op, b in ops[:3]:
        x, st = solve_alone(op, p, b, PREC.FP64)
        # assembled operator with identity rows on the constrained DoFs
        K = p.element_matrices()
        E, nd = p.gmap.shape
        rows = np.repeat(p.gmap.astype(np.int64), nd, axis=1).reshape(-1)
        cols = np.tile(p.gmap.astype(np.int64), (1, nd)).reshape(-1)
        A = sp.coo_matrix((K.reshape(-1), (rows, cols)), shape=(p.n_dofs, p.n_dofs)).tocsr()
        c = p.constrained.astype(np.int64)
        keep = np.ones(p.n_dofs, bool)
        keep[c] = False
        D = sp.diags(keep.astype(float))
        A = D @ A @ D + sp.diags((~keep).astype(float))
        ref = spla.spsolve(A.tocsc(), b)
        assert np.linalg.norm(x - ref) / np.linalg.norm(ref) <= 1e-8


def test_slots_budget_and_errors(ops):
    p, op, b = ops[0]
    s = Server(max_jobs=2)
    j1 = s.submit(op, PREC.FP64, b)
    j2 = s.submit(op, PREC.FP32, b)
    with pytest.raises(AmatError) as e:
        s.submit(op, PREC.BF16, b)
    assert e.value.name == "REJECTED"
    with pytest.raises(AmatError) as e:
        s.result(j1)  # not finished
    assert e.value.name == "INVALID_STATE"
    info = s.info()
    assert info.active_jobs == 2 and info.free_slots == 0
    assert info.workspace_bytes == Server.job_bytes(op, PREC.FP64) + Server.job_bytes(op, PREC.FP32)
    run_all(s, [j1, j2], {j1: None, j2: None})
    assert s.info().workspace_bytes == 0 and s.info().free_slots == 2
    with pytest.raises(AmatError) as e:
        s.result(j1)
    assert e.value.name == "INVALID_ARGUMENT"
    with pytest.raises(AmatError):
        s.submit(op, PREC.FP64, b, rtol=-1.0)


def test_serving_loop_small_stream(torch_gpu):
    """The policy loop end to end on a short stream of small requests: all
    converge, the budget is respected, the results equal solves alone."""
    mix = (("hex8", 0.6, 4, 9), ("hex20", 0.3, 2, 4), ("tet10", 0.1, 2, 4))
    reqs = serve.make_requests(24, rate=2000.0, seed=3, mix=mix)
    budget = 2 << 30
    for method in METHODS:
        cache = serve.OperatorCache()
        loop = serve.ServingLoop(Server(max_jobs=6), cache, method, budget, n_iter=4, keep_x=True)
        m = loop.run(reqs)
        bad = [(d.rid, d.kind, d.n, d.iterations, d.outer, d.rel_residual) for d in m["done"] if not d.converged]
        assert m["solves"] == len(reqs) and not bad, (method, bad)
        assert m["peak_resident_bytes"] <= budget and m["peak_jobs"] <= 6
        for d in m["done"][:6]:
            c = cache.get((d.kind, d.n))
            xa, _ = solve_alone(c.op, c.problem, rhs_for(c.problem, reqs[d.rid].seed), method)
            np.testing.assert_array_equal(d.x, xa)


def test_penalty_bc_records_served(torch_gpu):
    """hex20 elasticity with near-symmetric FP64 records and packed FP32 / BF16
    records under the penalty BC form (the record launch without the identity
    mask), beside an identity-BC job: served solves equal the single solves bit
    for bit."""
    from amat_amd import BCOP, Maps, aMatFree, problems
    torch = torch_gpu
    p = problems.elasticity_hex20(3, 3, 4)
    op = p.build_operator(aMatFree, Maps, precision=PREC.FP64 | PREC.FP32 | PREC.BF16, bcop=BCOP.PENALTY_DIAG,
                          penalty=1e8)
    assert op.info().near_symmetric == 1
    b = rhs_for(p, 17)
    other_p = build_problem("hex20", 3)
    other_op, other_b = build_operator(other_p), rhs_for(other_p, 18)
    # BF16 refinement under a 1e8 penalty: three outer steps of at most 400
    # iterations each (not converged), served and single alike
    for method, max_outer, max_it in ((PREC.FP64, 60, 20000), (PREC.FP32, 60, 20000), (FP32R, 60, 20000),
                                      (PREC.BF16, 3, 400)):
        db = torch.from_numpy(b).cuda()
        x = torch.zeros_like(db)
        if method == PREC.FP64:
            st = op.cg(db, x, rtol=1e-10, max_it=max_it)
        else:
            st = op.ir_cg(method, db, x, rtol=1e-10, inner_rtol=1e-2, max_outer=max_outer, max_inner=max_it)
        assert st.converged or method == PREC.BF16
        s = Server(max_jobs=4)
        j1 = s.submit(op, method, b, max_it=max_it, max_outer=max_outer)
        j2 = s.submit(other_op, method, other_b, max_it=max_it, max_outer=max_outer)
        res = run_all(s, [j1, j2], {j1: p.n_dofs, j2: other_p.n_dofs})
        xs, ss = res[j1]
        np.testing.assert_array_equal(xs, x.cpu().numpy())
        assert ss.iterations == st.iterations and ss.outer_iterations == st.outer_iterations
        if method != PREC.BF16:
            np.testing.assert_array_equal(res[j2][0], solve_alone(other_op, other_p, other_b, method)[0])
