"""Total-field / scattered-field plane-wave injection (evaluateTFSF,
HesthavenEvolution.cpp:97-124) in the oracle, pinned by physics rather than
by its formula: a Gaussian plane wave (E along z, propagating along +x)
injected through the plane x = 0.5 of a PEC(z)/PMC(y) waveguide with SMA
ends must reproduce the analytic incident wave on the total-field side and
leave the scattered-field side at rest.  A wrong sign or side puts the wave
on the wrong side or doubles it.  The product's host setup is checked to
classify the same face sides."""
import numpy as np
import pytest

import dgtd_amd as D
import meshgen as MG
from oracle import oracle as O

SPREAD, MEAN = 0.25, -0.6


def tfsf_plane(n):
    return 2.0 * max(1, n // 3) / n  # a grid plane of the stretched mesh


def waveguide(n):
    vx, E, _ = MG.cube_mesh(n, 1.0, "pec")
    vx = vx.copy()
    vx[:, 0] *= 2.0  # [0,2] x [0,1] x [0,1]
    bnd = MG.boundary_faces(E)
    fbc = np.zeros(E.shape, dtype=np.int32)
    for axis, code in ((0, 3), (1, 2), (2, 1)):  # x ends SMA, y walls PMC, z walls PEC
        for val in (0.0, 2.0 if axis == 0 else 1.0):
            fbc[MG.plane_faces(vx, E, axis, val) & bnd] = code
    assert (fbc[bnd] > 0).all()
    fbc[MG.plane_faces(vx, E, 0, tfsf_plane(n))] = 7
    assert (fbc == 7).sum() == 2 * 2 * n * n
    return vx, E, fbc


def incident(x, t):
    return np.exp(-((x - t) - MEAN) ** 2 / (2 * SPREAD ** 2))


def test_plane_wave_through_tfsf_plane():
    vx, E, fbc = waveguide(6)  # TF/SF plane x = 2/3
    ref = O.Operator(vx, E, fbc, 3)
    ref.set_planewaves([dict(pol=(0, 0, 1), dir=(1, 0, 0), spread=SPREAD, mean=MEAN)])
    x = np.zeros(ref.ndofs)
    T = 1.8  # the pulse peak crosses the plane at t = 1.27 and sits at x = 1.2 at T
    steps, t = ref.run(x, 0.5 / (ref.max_fscale() * 9), T, "lserk4")
    assert t == pytest.approx(T)
    pos = ref.positions().reshape(3, -1)
    f = x.reshape(6, -1)
    g = incident(pos[0], T)
    tf = pos[0] > 0.85
    sf = pos[0] < 0.5
    # total-field side: Ez = g, Hy = -g (H = k x pol = x x z = -y), the rest ~ 0
    assert np.linalg.norm(f[2][tf] - g[tf]) / np.linalg.norm(g[tf]) < 0.03
    assert np.linalg.norm(f[4][tf] + g[tf]) / np.linalg.norm(g[tf]) < 0.03
    for c in (0, 1, 3, 5):
        assert np.abs(f[c]).max() < 0.03
    # scattered-field side: nothing arrives
    assert np.abs(f[:, sf]).max() < 0.02


def test_no_wave_no_effect_and_linearity():
    vx, E, fbc = waveguide(2)
    plain = fbc.copy()
    plain[plain == 7] = 0
    rng = np.random.default_rng(0)
    a = O.Operator(vx, E, fbc, 2)
    b = O.Operator(vx, E, plain, 2)
    x = rng.standard_normal(a.ndofs)
    np.testing.assert_array_equal(a.mult(x, 0.3), b.mult(x, 0.3))  # no plane wave set
    a.set_planewaves([dict(pol=(0, 0, 1), dir=(1, 0, 0), spread=0.2, mean=0.1)])
    s = a.mult(np.zeros_like(x), 0.3)
    assert np.abs(s).max() > 1e-3
    np.testing.assert_allclose(a.mult(x, 0.3), b.mult(x) + s, rtol=0, atol=1e-12 * np.abs(s).max() + 1e-12)


def test_host_setup_classifies_tfsf_sides():
    vx, E, fbc = waveguide(2)
    M, P, codes = D.mesh_vmaps(vx, E, fbc.astype(np.int8), 2)
    plane = MG.plane_faces(vx, E, 0, tfsf_plane(2))
    # TF/SF faces stay ordinary interior faces for the jump (P = neighbour)
    Nfp = D.num_face_nodes(2)
    assert (P.reshape(-1, 4, Nfp)[plane] != M.reshape(-1, 4, Nfp)[plane]).all()
    np.testing.assert_array_equal(codes[plane], 7)
