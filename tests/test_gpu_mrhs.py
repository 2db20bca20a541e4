"""GPU parity of the multi-RHS MFMA products (amx_mrhs_*) against the oracle.

Every RHS column i of V = sum_e P_e^T K_e P_e U is checked against the
oracle's FP64 matvec of column i.  Two bars:
  * against the rounded model -- the oracle run with the same K and U the
    kernel multiplies (K rounded to FP32, or to BF16 after FP32, and for BF16
    also U rounded to BF16), so only FP32 accumulation differs: rel L2 <= 2e-6
    per column;
  * against the FP64 operator: FP32 storage <= 1e-6, BF16 storage <= 1e-2.
The atomic scatter (default) makes the FP32 sums order-dependent (tolerances
only); the deterministic scatter sums element rows per DoF in a fixed order:
its results are bit-identical run to run and across launch shapes.
"""
import numpy as np
import pytest

import oracle as orc
from amat_amd import PREC, SCATTER, AmatError, Maps, MultiRHS, aMatFree, mesh, problems
from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL_MODEL = 2e-6
# about twice the largest measured errors against the FP64 operator
TOL_FP64 = {PREC.FP32: 2e-7, PREC.BF16: 4e-3}


def bf16_round(a: np.ndarray) -> np.ndarray:
    """FP32 -> BF16 (round to nearest even) -> FP32, as the kernel rounds."""
    b = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def scalar_hex20(n: int, seed: int = 5) -> problems.Problem:
    """20 DoFs per element: one scalar per hex20 node, random NONsymmetric K_e
    (a transposed K_e in the kernel would fail the check)."""
    m = mesh.hex20(n, n, n, dofs_per_node=1)
    E = m.n_elem
    A = mesh.uniform(seed, 0, E * 400, -1.0, 1.0).reshape(E, 20, 20)
    K = A + 20.0 * np.eye(20)[None]
    return problems.Problem("scalar_hex20", m, 20, m.dof_map(), K=K)


CASES = {
    "hex20_elast": lambda: problems.elasticity_hex20(3, 4, 2),   # n = 60, 24 elements
    "hex8": lambda: problems.poisson_hex8(7, 5, 3),              # n = 8, 105 elements (ragged 4-element tiles)
    "tet10": lambda: problems.poisson_tet10(3, with_rhs=False),  # n = 10, 162 elements (3 per tile)
    "hex20_scalar": lambda: scalar_hex20(3),                     # n = 20
}


def make(torch, probs, precision):
    ops = [p.build_operator(aMatFree, Maps, precision=PREC.FP64) for p in probs]
    return ops, MultiRHS(ops, precision)


def inputs(torch, probs, nrhs, seed=11):
    U = [mesh.uniform(seed + i, 7, p.n_dofs * nrhs, -1.0, 1.0).astype(np.float32).reshape(p.n_dofs, nrhs)
         for i, p in enumerate(probs)]
    dU = [torch.from_numpy(u).cuda() for u in U]
    dV = [torch.full_like(u, float("nan")) for u in dU]
    return U, dU, dV


def check(p, precision, U, V):
    K64 = p.element_matrices()
    K32 = K64.astype(np.float32)
    Km = (bf16_round(K32) if precision == PREC.BF16 else K32).astype(np.float64)
    Um = (bf16_round(U) if precision == PREC.BF16 else U).astype(np.float64)
    worst_model = worst_64 = 0.0
    for i in range(U.shape[1]):
        ref_model = orc.matvec(p.gmap, Km, np.ascontiguousarray(Um[:, i]))
        ref_64 = orc.matvec(p.gmap, K64, np.ascontiguousarray(U[:, i].astype(np.float64)))
        worst_model = max(worst_model, rel_l2(V[:, i].astype(np.float64), ref_model))
        worst_64 = max(worst_64, rel_l2(V[:, i].astype(np.float64), ref_64))
    print(f"{p.name} {PREC(precision).name}: vs rounded model {worst_model:.2e}, vs FP64 {worst_64:.2e}")
    assert worst_model <= TOL_MODEL, (p.name, worst_model)
    assert worst_64 <= TOL_FP64[precision], (p.name, worst_64)
    return worst_model, worst_64


@pytest.mark.parametrize("scatter", [SCATTER.DETERMINISTIC, SCATTER.ATOMIC])
@pytest.mark.parametrize("precision", [PREC.FP32, PREC.BF16])
@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("nrhs", [1, 8, 37])
def test_single_operator(torch_gpu, case, precision, nrhs, scatter):
    torch = torch_gpu
    p = CASES[case]()
    ops, mr = make(torch, [p], precision)
    mr.set_scatter(scatter)
    U, dU, dV = inputs(torch, [p], nrhs)
    mr.apply_dev(nrhs, dU, dV, torch.cuda.current_stream())
    check(p, precision, U[0], dV[0].cpu().numpy())


@pytest.mark.parametrize("precision", [PREC.FP32, PREC.BF16])
def test_deterministic_bitwise_across_runs_and_shapes(torch_gpu, precision):
    torch = torch_gpu
    probs = [CASES["hex20_elast"](), CASES["hex8"](), CASES["tet10"]()]
    ops, mr = make(torch, probs, precision)
    mr.set_scatter(SCATTER.DETERMINISTIC)
    nrhs = 40
    U, dU, dV = inputs(torch, probs, nrhs)
    mr.apply_dev(nrhs, dU, dV, torch.cuda.current_stream())
    first = [v.cpu().numpy().copy() for v in dV]
    for rt, w, c in MultiRHS.TILE_SHAPES:
        mr.set_tiles(rt, w, c)
        for v in dV:
            v.fill_(float("nan"))
        mr.apply_dev(nrhs, dU, dV, torch.cuda.current_stream())
        for a, v in zip(first, dV):
            np.testing.assert_array_equal(v.cpu().numpy(), a)


@pytest.mark.parametrize("precision", [PREC.FP32, PREC.BF16])
def test_tile_shapes_agree(torch_gpu, precision):
    """Every (rhs_tile, waves_per_block, work_split) gives the same product within the bar."""
    torch = torch_gpu
    probs = [CASES["hex20_elast"](), CASES["hex8"]()]
    ops, mr = make(torch, probs, precision)
    nrhs = 70
    U, dU, dV = inputs(torch, probs, nrhs)
    for rt, w, c in MultiRHS.TILE_SHAPES:
        mr.set_tiles(rt, w, c)
        assert mr.tiles() == (rt, w, c)
        for v in dV:
            v.fill_(float("nan"))
        mr.apply_dev(nrhs, dU, dV, torch.cuda.current_stream())
        for p, u, v in zip(probs, U, dV):
            check(p, precision, u, v.cpu().numpy())


@pytest.mark.parametrize("precision", [PREC.FP32, PREC.BF16])
def test_persistent_mixed_meshes_autotune(torch_gpu, precision):
    """Four element types (60, 8, 10, 20 DoFs) in one persistent launch; the
    autotuner keeps a shape whose product still meets the bar."""
    torch = torch_gpu
    probs = [CASES[c]() for c in CASES]
    ops, mr = make(torch, probs, precision)
    nrhs = 32
    U, dU, dV = inputs(torch, probs, nrhs)
    times = mr.autotune(nrhs, dU, dV, reps=3, stream=torch.cuda.current_stream())
    assert times and all(t > 0 for t in times.values())
    assert mr.tiles() in times
    for v in dV:
        v.fill_(float("nan"))
    mr.apply_dev(nrhs, dU, dV, torch.cuda.current_stream())
    for p, u, v in zip(probs, U, dV):
        check(p, precision, u, v.cpu().numpy())


def test_graph_replays(torch_gpu):
    """A captured apply replayed several times, with eager work and new U
    contents between replays, matches the eager apply each time."""
    torch = torch_gpu
    probs = [CASES["hex20_elast"](), CASES["tet10"]()]
    ops, mr = make(torch, probs, PREC.FP32)
    nrhs = 16
    U, dU, dV = inputs(torch, probs, nrhs)
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    mr.graph_capture(nrhs, dU, dV, s)
    eV = [torch.empty_like(v) for v in dV]
    for rep in range(4):
        with torch.cuda.stream(s):
            for i, u in enumerate(dU):
                u.copy_(torch.from_numpy(U[i] * (rep + 1)).cuda())
            for v in dV:
                v.fill_(float("nan"))
        mr.graph_launch(s)
        mr.apply_dev(nrhs, dU, eV, s)
        s.synchronize()
        for p, u, v, e in zip(probs, U, dV, eV):
            vn = v.cpu().numpy()
            assert rel_l2(vn.astype(np.float64), e.cpu().numpy().astype(np.float64)) <= 1e-6
            check(p, PREC.FP32, u * (rep + 1), vn)


def test_errors(torch_gpu):
    p = CASES["hex8"]()
    op = p.build_operator(aMatFree, Maps)
    with pytest.raises(AmatError) as e:
        MultiRHS([op], PREC.FP64)
    assert e.value.name == "INVALID_ARGUMENT"
    m = mesh.hex8(2, 2, 2)
    odd = problems.Problem("n12", m, 12, np.concatenate([m.dof_map(), m.dof_map()[:, :4]], axis=1),
                           K=np.tile(np.eye(12), (m.n_elem, 1, 1)))
    op12 = odd.build_operator(aMatFree, Maps)
    with pytest.raises(AmatError) as e:
        MultiRHS([op12], PREC.FP32)
    assert e.value.name == "NOT_IMPLEMENTED"
    mr = MultiRHS([op], PREC.FP32)
    with pytest.raises(AmatError):
        mr.set_tiles(48, 1)
    with pytest.raises(AmatError):
        mr.set_tiles(32, 1, 2)
    with pytest.raises(AmatError):
        mr.graph_launch(None)
