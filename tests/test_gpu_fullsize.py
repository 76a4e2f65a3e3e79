"""Checks at the benchmark sizes, through the C ABI.

Workloads: the PEC cavity [0,1]^3 with 32^3 hexahedra x 6 = 196,608 tets at
order 4 (41,287,680 DOF), and the SMA box [0,2]^3 with 48^3 x 6 = 663,552
tets at order 3 with eps_r ~ U[1,4] (79,626,240 DOF).

- FP64 Mult equals the oracle's Mult on the same seeded input to 1e-12
  relative (the oracle runs one application, multi-threaded on the host);
- the two matrix-free variants (per-node VALU and MFMA) agree;
- FP32 agrees with FP64 to 1e-5 relative L2;
- Mult is linear;
- the energy identities of the semi-discrete operator hold at full size: with
  the centred flux and PEC walls the operator is skew in the energy inner
  product, B(x, A x) = 0; the upwind flux dissipates, B(x, A x) < 0.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

import dgtd_amd as D  # noqa: E402


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def maxrel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.fixture(scope="module", autouse=True)
def oracle_threads():
    O.set_threads(max(1, min(16, os.cpu_count() or 1)))


@pytest.fixture(scope="module")
def cavity32():
    vx, E, fbc = D.cube_mesh(32, 1.0, "pec")
    x = np.random.default_rng(32).standard_normal(6 * D.num_nodes(4) * (E.size // 4))
    return vx, E, fbc, x


def test_cavity32_p4_mult_matches_oracle_both_variants(cavity32):
    vx, E, fbc, x = cavity32
    yr = O.Operator(vx, E, fbc, 4).mult(x)
    ys = {}
    for kernel in ("valu", "mfma"):
        g = D.HesthavenEvolution(vx, E, fbc, 4, precision=64, kernel=kernel)
        ys[kernel] = g.mult(x)
        g.close()
        assert maxrel(ys[kernel], yr) <= 1e-12, kernel
    assert maxrel(ys["mfma"], ys["valu"]) <= 1e-12
    g32 = D.HesthavenEvolution(vx, E, fbc, 4, precision=32, kernel="mfma")
    assert rel(g32.mult(x), yr) <= 1e-5
    g32.close()


def test_cavity32_p4_linearity(cavity32):
    vx, E, fbc, x = cavity32
    y = np.random.default_rng(7).standard_normal(x.size)
    g = D.HesthavenEvolution(vx, E, fbc, 4, precision=64, kernel="mfma")
    lhs = g.mult(2.5 * x - 0.75 * y)
    rhs = 2.5 * g.mult(x) - 0.75 * g.mult(y)
    assert maxrel(lhs, rhs) <= 1e-12
    g.close()


@pytest.mark.parametrize("alpha", [0.0, 1.0])
def test_cavity32_p4_energy_identities(cavity32, alpha):
    vx, E, fbc, x = cavity32
    g = D.HesthavenEvolution(vx, E, fbc, 4, alpha=alpha, precision=64, kernel="mfma")
    y = g.mult(x)

    def energy(v):
        g.set_state(v)
        return g.norms()["energy"]

    b = 0.5 * (energy(x + y) - energy(x - y))  # B(x, A x)
    scale = np.sqrt(energy(x) * energy(y))
    if alpha == 0.0:
        assert abs(b) <= 1e-11 * scale
    else:
        assert b < -1e-3 * scale
    g.close()


def test_sma_box48_p3_mult_matches_oracle():
    vx, E, fbc = D.cube_mesh(48, 2.0, "sma")
    K = E.size // 4
    eps = np.random.default_rng(2024).uniform(1.0, 4.0, K)
    x = np.random.default_rng(48).standard_normal(6 * D.num_nodes(3) * K)
    yr = O.Operator(vx, E, fbc, 3, eps=eps).mult(x)
    g = D.HesthavenEvolution(vx, E, fbc, 3, eps=eps, precision=64)
    assert maxrel(g.mult(x), yr) <= 1e-12
    g.close()
    g32 = D.HesthavenEvolution(vx, E, fbc, 3, eps=eps, precision=32)
    assert rel(g32.mult(x), yr) <= 1e-5
    g32.close()
