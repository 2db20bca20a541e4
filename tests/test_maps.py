"""DoF maps, BC maps and element lists: the oracle restatement and the
product's Maps (C ABI) against fixtures written by the reference
include/maps.hpp run under MPICH at 1-3 ranks (tests/golden/make_golden.py).
Integer outputs must be bit-exact."""
import numpy as np
import pytest

import oracle as orc
from amat_amd import AmatError, Maps
from conftest import GOLDEN

CASES = sorted(p.stem for p in GOLDEN.glob("maps_*.npz"))


def load(case):
    z = np.load(GOLDEN / f"{case}.npz")
    return {k: z[k] for k in z.files}


def alltoallv(requests):
    """Simulate the Alltoallv of buildScatterMap: requests[r] = (counts per
    owner, gids grouped by owner) -> what each owner receives, grouped by
    requesting rank."""
    P = len(requests)
    got = []
    for p in range(P):
        counts, chunks = np.zeros(P, np.uint32), []
        for r in range(P):
            c, g = requests[r]
            off = int(np.sum(c[:p]))
            counts[r] = c[p]
            chunks.append(g[off:off + int(c[p])])
        got.append((counts, np.concatenate(chunks).astype(np.uint64) if chunks else np.zeros(0, np.uint64)))
    return got


def test_fixtures_present():
    assert len(CASES) >= 8


@pytest.mark.parametrize("case", CASES)
def test_product_maps_match_reference(case):
    z = load(case)
    P = int(z["size"][0])
    counts = z["rank_dof_counts"]
    maps, reqs = [], []
    for r in range(P):
        m = Maps(r, P)
        e2r = z[f"r{r}_e2r"]
        b, e, ng = (int(x) for x in z[f"r{r}_range"])
        m.set_map(e2r.shape[0], e2r, None, len(z[f"r{r}_r2g"]), z[f"r{r}_r2g"], b, e, ng, counts)
        reqs.append(m.get_ghost_requests())
        maps.append(m)
    for r, (c, g) in enumerate(alltoallv(reqs)):
        maps[r].set_peer_requests(c, g)
    for r in range(P):
        m = maps[r]
        m.set_bdr_map(z[f"r{r}_cd"], z[f"r{r}_cv"])
        ref = lambda k: z[f"r{r}_ref_{k}"]  # noqa: E731
        cnt = z[f"r{r}_counts"]
        assert (m.get_NumPreGhostDofs(), m.get_NumDofs(), m.get_NumPostGhostDofs(), m.get_NumDofsTotal()) == tuple(cnt)
        np.testing.assert_array_equal(m.get_LocalMap(), ref("local_map"))
        np.testing.assert_array_equal(m.get_Map(), ref("global_map"))
        # local-to-global entries referenced by elements (others are unset in the reference)
        used = np.unique(ref("local_map"))
        np.testing.assert_array_equal(m.get_Local2Global()[used], ref("local2global")[used])
        np.testing.assert_array_equal(m.get_localID2Rank(), ref("localid2rank"))
        np.testing.assert_array_equal(m.get_SendDofCounts(), ref("send_counts"))
        np.testing.assert_array_equal(m.get_SendDofOffset(), ref("send_offsets"))
        np.testing.assert_array_equal(m.get_SendDofIds(), ref("send_ids"))
        np.testing.assert_array_equal(m.get_SendRankIds(), ref("send_rank_ids"))
        np.testing.assert_array_equal(m.get_RecvDofCounts(), ref("recv_counts"))
        np.testing.assert_array_equal(m.get_RecvDofOffset(), ref("recv_offsets"))
        np.testing.assert_array_equal(m.get_RecvRankIds(), ref("recv_rank_ids"))
        np.testing.assert_array_equal(m.get_independentElem(), ref("independent"))
        np.testing.assert_array_equal(m.get_dependentElem(), ref("dependent"))
        np.testing.assert_array_equal(m.get_BdrMap(), ref("bdr_map"))
        np.testing.assert_array_equal(m.get_PresValMap(), ref("presval"))
        np.testing.assert_array_equal(m.get_ownedConstrainedDofs(), ref("owned_cdofs"))
        np.testing.assert_array_equal(m.get_ownedPrescribedValues(), ref("owned_cvals"))
        np.testing.assert_array_equal(m.get_ownedFreeDofs(), ref("owned_free"))
        np.testing.assert_array_equal(m.get_LocalDofScan(), ref("dof_scan"))


@pytest.mark.parametrize("case", CASES)
def test_oracle_maps_match_reference(case):
    """The oracle's restatement against the same fixtures (pins the oracle)."""
    z = load(case)
    P = int(z["size"][0])
    counts = z["rank_dof_counts"]
    outs, reqs = [], []
    for r in range(P):
        e2r = z[f"r{r}_e2r"]
        b, e, ng = (int(x) for x in z[f"r{r}_range"])
        o = orc.build_scatter_map(r, P, counts, e2r, len(z[f"r{r}_r2g"]), z[f"r{r}_r2g"], b, e, ng)
        assert o["err"] == 0
        outs.append(o)
        reqs.append((o["ghost_req_counts"], o["ghost_req_gids"]))
    for r, (c, g) in enumerate(alltoallv(reqs)):
        ref = lambda k: z[f"r{r}_ref_{k}"]  # noqa: E731
        o = outs[r]
        np.testing.assert_array_equal(o["local_map"], ref("local_map"))
        np.testing.assert_array_equal(o["global_map"], ref("global_map"))
        used = np.unique(ref("local_map"))
        np.testing.assert_array_equal(o["local2global"][used], ref("local2global")[used])
        np.testing.assert_array_equal(o["localid2rank"], ref("localid2rank"))
        np.testing.assert_array_equal(c, ref("send_counts"))
        err, ids = orc.set_send_ids(r, P, counts, o["n_pre"], g)
        assert err == 0
        np.testing.assert_array_equal(ids, ref("send_ids"))
        np.testing.assert_array_equal(o["ghost_req_counts"], ref("recv_counts"))
        gm2 = o["global_map"].reshape(z[f"r{r}_e2r"].shape)
        bd = orc.set_bdr_map(r, P, counts, gm2, z[f"r{r}_cd"], z[f"r{r}_cv"])
        np.testing.assert_array_equal(bd["bdr_map"], ref("bdr_map"))
        np.testing.assert_array_equal(bd["presval_map"], ref("presval"))
        np.testing.assert_array_equal(bd["owned_cdofs"], ref("owned_cdofs"))
        np.testing.assert_array_equal(bd["owned_cvals"], ref("owned_cvals"))
        np.testing.assert_array_equal(bd["owned_free"], ref("owned_free"))
        ind, dep = orc.identify_independent(r, P, counts, gm2)
        np.testing.assert_array_equal(ind, ref("independent"))
        np.testing.assert_array_equal(dep, ref("dependent"))


def test_one_rank_local_map_is_caller_map():
    """With one rank the ghost lists are empty and the local map is the
    caller's map (the situation of example03a/07a at one rank)."""
    from amat_amd import mesh
    me = mesh.hex20(2, 3, 2)
    g = me.dof_map().astype(np.uint32)
    m = Maps()
    n = me.n_dofs
    m.set_map(me.n_elem, g, None, n, np.arange(n), 0, n - 1, n)
    np.testing.assert_array_equal(m.get_LocalMap(), g.reshape(-1))
    assert m.get_NumPreGhostDofs() == 0 and m.get_NumPostGhostDofs() == 0
    assert len(m.get_dependentElem()) == 0


def test_bdr_map_first_occurrence_wins_and_scales():
    """Duplicate constraint ids keep the first value (the reference's break);
    the hash path must agree with a literal scan on a larger map."""
    from amat_amd import mesh
    me = mesh.hex8(6, 5, 4)
    g = me.dof_map()
    n = me.n_dofs
    rng = np.random.default_rng(3)
    cd = rng.choice(n, 60, replace=False).astype(np.uint64)
    cd = np.concatenate([cd, cd[:10]])
    cv = rng.uniform(-1, 1, len(cd))
    m = Maps()
    m.set_map(me.n_elem, g.astype(np.uint32), None, n, np.arange(n), 0, n - 1, n)
    m.set_bdr_map(cd, cv)
    o = orc.set_bdr_map(0, 1, [n], g, cd, cv)
    np.testing.assert_array_equal(m.get_BdrMap(), o["bdr_map"])
    np.testing.assert_array_equal(m.get_PresValMap(), o["presval_map"])
    np.testing.assert_array_equal(m.get_ownedConstrainedDofs(), o["owned_cdofs"])
    np.testing.assert_array_equal(m.get_ownedFreeDofs(), o["owned_free"])


def test_error_codes():
    m = Maps(0, 2)
    e2r = np.array([[0, 1, 2, 3]], np.uint32)
    # a ghost id beyond every rank's range
    with pytest.raises(AmatError) as ei:
        m.set_map(1, e2r, None, 4, np.array([0, 1, 2, 99]), 0, 2, 6, [3, 3])
    assert ei.value.name == "GHOST_NODE_NOT_FOUND"
    with pytest.raises(AmatError) as ei:
        m.set_map(1, e2r, None, 5, np.array([0, 1, 2, 3, 4]), 0, 2, 6, [3, 3])
    assert ei.value.name == "INVALID_ARGUMENT"   # n_all_dofs disagrees with owned + ghosts
    with pytest.raises(AmatError) as ei:
        m.set_map(1, np.array([[0, 1, 2, 7]], np.uint32), None, 4, np.array([0, 1, 2, 3]), 0, 2, 6, [3, 3])
    assert ei.value.name == "INDEX_OUT_OF_BOUNDS"
    with pytest.raises(AmatError) as ei:
        Maps().set_bdr_map([0], [0.0])
    assert ei.value.name == "INVALID_STATE"
    m.set_map(1, e2r, None, 4, np.array([0, 1, 2, 3]), 0, 2, 6, [3, 3])
    with pytest.raises(AmatError) as ei:
        m.set_peer_requests([0, 1], [5])   # a peer asks for an id this rank does not own
    assert ei.value.name == "GHOST_NODE_NOT_FOUND"
