"""Element matrices (amx_fem_*) against the reference's own ke_matrix.cpp /
fe_vector.cpp (fixtures) and against what the mathematics fixes: the closed
form of the hex8 Laplacian on a cube, symmetry, null spaces (constants,
rigid-body modes), exactness of the Gauss rule, volumes."""
import numpy as np
import pytest

from amat_amd import fem, mesh
from conftest import GOLDEN


@pytest.fixture(scope="module")
def ref():
    z = np.load(GOLDEN / "fem_ref.npz")
    return {k: z[k] for k in z.files}


def close(a, b, rtol=1e-13):
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() <= rtol * scale


def test_gauss_matches_reference_and_numpy(ref):
    for n in range(1, 6):
        g = fem.gauss(n)
        np.testing.assert_allclose(g, ref[f"gauss_{n}"], rtol=0, atol=1e-15)
        x, w = np.polynomial.legendre.leggauss(n)
        np.testing.assert_allclose(g[0::2], x, atol=1e-14)
        np.testing.assert_allclose(g[1::2], w, atol=1e-14)


def test_ke_hex8_matches_reference(ref):
    for key in [k for k in ref if k.startswith("hex8_") and k.endswith("_ke")]:
        xe = ref[key[:-3] + "_xe"]
        ngt = int(key.split("_ngt")[1][0])
        assert close(fem.ke_hex8(xe, ngt), ref[key]), key
    for name in ("cube32", "cube1", "distorted"):
        assert close(fem.ke_hex8(ref[f"hex8_{name}_ngt2_xe"], 2), ref[f"hex8eig_{name}_ke"])


def test_ke_hex8_cube_closed_form():
    """Q1 Laplacian on a cube of side h: K = (h/12) A with A_ii = 4, 0 for edge
    neighbours, -1 for face and body diagonals (2x2x2 Gauss is exact)."""
    from amat_amd.problems import hex8_cube_ke
    corners = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
    A = np.zeros((8, 8))
    for i in range(8):
        for j in range(8):
            d = np.abs(corners[i] - corners[j]).sum()
            A[i, j] = 4.0 if d == 0 else (0.0 if d == 1 else -1.0)
    for h in (1.0, 1 / 32, 0.37):
        np.testing.assert_allclose(hex8_cube_ke(h), (h / 12.0) * A, rtol=1e-13, atol=1e-15 * h)


def test_ke_hex20_matches_reference(ref):
    for key in [k for k in ref if k.startswith("hex20_") and k.endswith("_ke")]:
        xe = ref[key[:-3] + "_xe"]
        parts = key.split("_")
        E = float(parts[2][1:])
        nu = float(parts[3][2:])
        ngt = int(parts[4][3:])
        assert close(fem.ke_hex20_iso(xe, E, nu, ngt), ref[key]), key


def rigid_modes(xe):
    modes = []
    for c in range(3):
        m = np.zeros((20, 3))
        m[:, c] = 1.0
        modes.append(m.reshape(-1))
    for (a, b) in ((0, 1), (1, 2), (0, 2)):
        m = np.zeros((20, 3))
        m[:, a] = -xe[:, b]
        m[:, b] = xe[:, a]
        modes.append(m.reshape(-1))
    return np.stack(modes, axis=1)


@pytest.mark.parametrize("ngt", [2, 3])
def test_ke_hex20_invariants(ngt):
    me = mesh.hex20(1, 1, 1)
    rng = np.random.default_rng(11)
    for xe in (me.coords[me.elem_nodes[0]] / 64.0, me.coords[me.elem_nodes[0]] + rng.uniform(-0.08, 0.08, (20, 3))):
        K = fem.ke_hex20_iso(xe, 1.0, 0.3, ngt)
        assert np.abs(K - K.T).max() <= 1e-13 * np.abs(K).max()
        R = rigid_modes(xe)
        assert np.abs(K @ R).max() <= 1e-11 * np.abs(K).max() * np.abs(R).max()
        ev = np.linalg.eigvalsh(K)
        if ngt == 3:  # full integration: exactly the six rigid-body zero modes
            assert (ev[6:] > 1e-8 * ev[-1]).all()   # (2x2x2 reduced integration has spurious ones)


def test_ke_hex20_gauss_exact_on_affine_cube():
    """On an affine cube the hex20 integrand is a polynomial of degree <= 4 per
    direction: 3x3x3 Gauss equals 4x4x4."""
    me = mesh.hex20(1, 1, 1)
    xe = me.coords[me.elem_nodes[0]] / 64.0
    K3 = fem.ke_hex20_iso(xe, 1.0, 0.3, 3)
    K4 = fem.ke_hex20_iso(xe, 1.0, 0.3, 4)
    np.testing.assert_allclose(K3, K4, rtol=0, atol=1e-13 * np.abs(K4).max())


def test_ke_tet10_matches_reference(ref):
    for key in [k for k in ref if k.startswith("tet10_") and k.endswith("_ke")]:
        base = key[:-3]
        ngt = int(base.split("ngt")[1])
        xe = ref[base + "_xe"]
        assert close(fem.ke_tet10(xe, ngt), ref[key]), key
        assert close(fem.fe_tet10(xe, ref[base + "_be"], ngt), ref[base + "_fe"]), key


def test_tet10_invariants_and_batch():
    m = mesh.tet10(3, seed=9, jitter=0.2)
    coords = m.coords[m.elem_nodes]
    K = fem.ke_tet10_batch(coords, 2)
    assert np.abs(K.sum(axis=2)).max() <= 1e-12 * np.abs(K).max()   # constants in the null space
    assert np.abs(K - K.transpose(0, 2, 1)).max() <= 1e-13 * np.abs(K).max()
    for e in (0, 17, len(K) - 1):
        np.testing.assert_allclose(K[e], fem.ke_tet10(coords[e], 2), rtol=0, atol=1e-15 * np.abs(K[e]).max())
    # consistent-mass load with f = 1 integrates the volume; volumes sum to 1
    fe = fem.fe_tet10_batch(coords, np.ones((len(K), 10)), 2)
    v = coords[:, :4]
    vol = np.abs(np.linalg.det(np.stack([v[:, 1] - v[:, 0], v[:, 2] - v[:, 0], v[:, 3] - v[:, 0]], axis=1))) / 6
    np.testing.assert_allclose(fe.sum(axis=1), vol, rtol=1e-12)
    assert abs(vol.sum() - 1.0) < 1e-12


def test_negative_jacobian_is_an_error():
    from amat_amd import AmatError
    xe = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    with pytest.raises(AmatError):
        fem.ke_hex8(xe[[1, 0, 3, 2, 5, 4, 7, 6]], 2)   # mirrored element
