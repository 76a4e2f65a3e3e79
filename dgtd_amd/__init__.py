"""dgtd_amd -- MI355X-native (gfx950) hesthaven DGTD Maxwell operator.

The compute path is libdgtd_hip.so (hand-written HIP kernels behind the C ABI
in include/dgtd_hip.h); this package is the host-side mirror of the
reference's operator and solver surfaces.
"""
from ._abi import DgtdError, LIB_PATH, lib  # noqa: F401
from .batch import Batch, CapacityError, footprint  # noqa: F401
from .hesthaven import (CANDIDATES, HesthavenEvolution, GlobalEvolution, Solver, cube_mesh, square_mesh,  # noqa: F401
                        bessel_j6, cartesian_2d, cartesian_3d, mesh_time_step, estimate_time_step, gaussian, global_assemble, mesh_geometry, mesh_locate, mesh_vmaps, num_face_nodes, num_nodes,
                        planewave, project_initial_fields, ref_element, resonant)
from .case import Case, run_case  # noqa: F401,E402
