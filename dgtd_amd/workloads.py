"""Synthetic workloads of the benchmark configurations, shared by bench.py and
the tests (no counterpart in the reference, which runs one case file per
process).

cavity_batch: independent PEC cavities [0,1]^3 of n^3 hexahedra x 6
tetrahedra, n log-uniform over {2..24}, order uniform over {1..5}, each with a
seeded Gaussian E_z pulse (width 0.1, centre uniform in [0.3,0.7]^3) -- the
batched-instances configuration.  The draws per instance are n, order, then
the centre, in that order, from numpy's default_rng(seed).
"""
import numpy as np

from .hesthaven import cube_mesh, gaussian, num_nodes, project_initial_fields


def cavity_batch_params(seed=2024, count=64):
    """[(n, order, centre)] of the batch, without building anything."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(round(np.exp(rng.uniform(np.log(2), np.log(24)))))
        order = int(rng.integers(1, 6))
        c = rng.uniform(0.3, 0.7, 3)
        out.append((n, order, c))
    return out


def cavity_instance(n, order, centre):
    """(vx, EToV, face_bc, x0) of one cavity of the batch."""
    vx, E, fbc = cube_mesh(n, 1.0, "pec")
    x0 = project_initial_fields(vx, E, order, [gaussian("E", centre, 0.1, (0, 0, 1))])
    return vx, E, fbc, x0


def cavity_batch(seed=2024, count=64):
    """[(n, order, vx, EToV, face_bc, x0)] for every instance of the batch."""
    return [(n, o) + cavity_instance(n, o, c) for (n, o, c) in cavity_batch_params(seed, count)]


def instance_dof(n, order):
    return 6 * num_nodes(order) * 6 * n ** 3
