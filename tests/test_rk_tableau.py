"""The oracle's time integrators pinned by their stability polynomials.

For a linear operator L one step of an explicit Runge-Kutta method is
R(dt L) q with R(z) = sum_k gamma_k z^k.  Fourth order on linear problems
means gamma_k = 1/k! for k <= 4; classic RK4 (mfem::RK4Solver, the
reference's integrator, Solver.cpp:45-47) has R(z) = sum_{k<=4} z^k/k!
exactly, and the five-stage 2N-storage LSERK4 of Carpenter & Kennedy (1994)
has one more term gamma_5 z^5.  This file transcribes the published LSERK4
coefficients once more, checks them against the order conditions through
the scheme's own recurrence (a wrong digit moves some gamma_k, k <= 4), and
checks the oracle's steps against R(dt L) q built from its Mult to 1e-14 --
so a typo in the oracle's copy of the tableau fails here.  The stage times
c_s (used by the TF/SF sources) are pinned as the z^1 coefficient of each
stage's input, and the oracle's TF/SF step against a stage-by-stage step
from Mult at those times."""
import math

import numpy as np
import pytest

import meshgen as MG
from oracle import oracle as O

# Carpenter & Kennedy (1994), NASA TM-109112, five-stage fourth-order
# 2N-storage scheme: res = A_s res + dt L(q, t + C_s dt); q = q + B_s res
A = [0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
     -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0]
B = [1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
     1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
     2277821191437.0 / 14882151754819.0]


def lserk_polynomials():
    """Stability polynomial of the 2N-storage recurrence on y' = z y (as
    coefficient arrays in z), and the z^1 coefficient of each stage input."""
    q = np.zeros(7)
    q[0] = 1.0
    res = np.zeros(7)
    c = []
    for s in range(5):
        c.append(q[1])
        zq = np.roll(q, 1)
        zq[0] = 0.0
        res = A[s] * res + zq
        q = q + B[s] * res
    return q, c


def test_lserk4_tableau_meets_the_order_conditions():
    g, c = lserk_polynomials()
    for k in range(5):
        assert g[k] == pytest.approx(1.0 / math.factorial(k), abs=1e-15)
    assert 0.0 < g[5] < 1.0 / 120.0 and g[6] == 0.0  # five stages: degree 5
    assert c[0] == 0.0 and all(0.0 < x < 1.0 for x in c[1:])


def small_operator(order=2):
    vx, E, fbc = MG.cube_mesh(2, 1.0, ["pec", "pmc", "sma", "pec", "pmc", "sma"])
    rng = np.random.default_rng(5)
    K = E.shape[0]
    ref = O.Operator(vx, E, fbc, order, eps=rng.uniform(1, 3, K), mu=rng.uniform(1, 2, K))
    q = rng.standard_normal(ref.ndofs)
    dt = 1.0 / (ref.max_fscale() * order ** 2)
    return ref, q, dt


def powers(ref, q, dt, n):
    out = [q.copy()]
    for _ in range(n):
        out.append(dt * ref.mult(out[-1]))
    return out


def test_oracle_lserk4_step_is_its_stability_polynomial():
    ref, q, dt = small_operator()
    g, _ = lserk_polynomials()
    P = powers(ref, q, dt, 5)
    want = sum(g[k] * P[k] for k in range(6))
    x, res = q.copy(), np.zeros_like(q)
    ref.lserk4_step(x, res, 0.0, dt)
    assert np.abs(x - want).max() <= 1e-14 * np.abs(want).max()
    # and gamma_5 matters: the 4th-order truncation is further off
    trunc = sum(P[k] / math.factorial(k) for k in range(5))
    assert np.abs(x - trunc).max() > 1e3 * np.abs(x - want).max()


def test_oracle_rk4_step_is_the_taylor_polynomial():
    ref, q, dt = small_operator()
    P = powers(ref, q, dt, 4)
    want = sum(P[k] / math.factorial(k) for k in range(5))
    x = q.copy()
    ref.rk4_step(x, 0.0, dt)
    assert np.abs(x - want).max() <= 1e-14 * np.abs(want).max()


def test_oracle_lserk4_stage_times_with_a_time_dependent_source():
    """TF/SF plane waves make Mult affine and time dependent: the oracle's
    step equals the recurrence run on Mult at t + c_s dt."""
    vx, E, fbc = MG.cube_mesh(3, 1.0, "pec")
    fbc = np.asarray(fbc, dtype=np.int32)
    fbc[MG.plane_faces(vx, E, 0, 1.0 / 3.0)] = 7
    ref = O.Operator(vx, E, fbc, 2)
    ref.set_planewaves([dict(pol=(0, 0, 1), dir=(1, 0, 0), spread=0.1, mean=0.0)])
    rng = np.random.default_rng(3)
    q0 = 0.1 * rng.standard_normal(ref.ndofs)
    dt, t0 = 0.5 / (ref.max_fscale() * 4), 0.2
    _, c = lserk_polynomials()
    q, res = q0.copy(), np.zeros_like(q0)
    for s in range(5):
        res = A[s] * res + dt * ref.mult(q, t0 + c[s] * dt)
        q = q + B[s] * res
    x, r = q0.copy(), np.zeros_like(q0)
    ref.lserk4_step(x, r, t0, dt)
    assert np.abs(x - q).max() <= 1e-14 * np.abs(q).max()
    # the source does depend on the stage times
    q2, res2 = q0.copy(), np.zeros_like(q0)
    for s in range(5):
        res2 = A[s] * res2 + dt * ref.mult(q2, t0)
        q2 = q2 + B[s] * res2
    assert np.abs(q2 - q).max() > 1e-6 * np.abs(q).max()
