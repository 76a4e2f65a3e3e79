/*
 * dgtd_hip.h -- C ABI of the MI355X (gfx950) implementation of OpenSEMBA/dgtd's
 * "hesthaven" DGTD Maxwell evolution operator and its Runge-Kutta stage update.
 *
 * The library replaces, for one GPU:
 *   - maxwell::HesthavenEvolution (src/evolution/HesthavenEvolution.h:14-50),
 *     its constructor (HesthavenEvolution.cpp:315-437) and
 *     Mult(const mfem::Vector&, mfem::Vector&) const (HesthavenEvolution.cpp:450-542);
 *   - the RK stage arithmetic of mfem::RK4Solver as driven by Solver::step /
 *     Solver::run (src/solver/Solver.cpp:483-551, RK4 chosen at :45-47);
 *   - the Hesthaven 3D time-step rule (Solver.cpp:334-346), via max fscale.
 *
 * Conventions
 *   - State vectors use the reference layout (src/evolution/Fields.h:48-56):
 *     x[c*N + e*Np + i], c in {Ex,Ey,Ez,Hx,Hy,Hz}, N = Np*K, Np = (p+1)(p+2)(p+3)/6.
 *   - Host entry points take host pointers and FP64 values whatever the
 *     operator's precision; *_device entry points take device pointers in the
 *     operator's precision (double for 64, float for 32).
 *   - Every function returns 0 on success or a negative DGTD_E* code; the
 *     message of the last failure on the calling thread is dgtd_last_error().
 *     No C++ exception crosses this boundary.
 *   - A handle owns its device buffers and one HIP stream.  Calls on one handle
 *     must be serialised by the caller; different handles are independent.
 *     The library keeps no caller pointer after a call returns.
 */
#ifndef DGTD_HIP_H
#define DGTD_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DGTD_OK 0
#define DGTD_E_ARG (-1)         /* bad argument */
#define DGTD_E_NOMEM (-2)       /* device or host allocation failed */
#define DGTD_E_HIP (-3)         /* HIP runtime error */
#define DGTD_E_UNSUPPORTED (-4) /* feature outside this implementation */
#define DGTD_E_STATE (-5)       /* call not valid in the handle's state */
#define DGTD_E_CAPACITY (-6)    /* admission refused: HBM budget or slots */

/* Face boundary codes (face f of a tetrahedron is opposite local vertex f,
 * as MFEM's tetrahedron face table orders them).  Mirrors BdrCond PEC/PMC/SMA
 * (src/components/Types.h:47-55) as applied in
 * HesthavenEvolution::applyBoundaryConditionsToNodes (HesthavenEvolution.cpp:275-313). */
#define DGTD_BC_NONE 0 /* interior face, or boundary face with zero jump */
#define DGTD_BC_PEC 1
#define DGTD_BC_PMC 2
#define DGTD_BC_SMA 3
/* Interior boundary conditions on interior faces (given on either side; the
 * jump on both sides becomes (-1, 0), (0, -1), (-1/2, -1/2) times the own
 * trace, HesthavenEvolution.cpp:294-311). */
#define DGTD_BC_INT_PEC 4
#define DGTD_BC_INT_PMC 5
#define DGTD_BC_INT_SMA 6
/* Total-field / scattered-field surface (interior faces, either side): the
 * plane waves of dgtd_op_set_planewaves are injected there
 * (HesthavenEvolution::evaluateTFSF, HesthavenEvolution.cpp:97-124).  The
 * face's lower-index element is the scattered-field side, the face's first
 * element as MFEM orders them (HesthavenEvolutionMethods.cpp:562-574). */
#define DGTD_BC_TFSF 7

/* Integrators. */
#define DGTD_RK4 0    /* mfem::RK4Solver stage order (the reference's default) */
#define DGTD_LSERK4 1 /* Carpenter-Kennedy 5-stage low-storage RK4 */

typedef struct dgtd_op dgtd_op;

/* Mesh description; host arrays borrowed for the duration of dgtd_op_create.
 * Replaces the (ParFiniteElementSpace, Model) pair of the reference
 * constructor (HesthavenEvolution.cpp:315): affine tetrahedra, or affine
 * triangles with dim = 2; quadrilaterals and hexahedra per nvert. */
typedef struct {
    int32_t K;              /* number of elements */
    int32_t order;          /* polynomial order p, 1..5 */
    const double* vx;       /* 3*nv vertex coordinates (x,y,z interleaved) */
    int32_t nv;             /* number of vertices */
    const int32_t* EToV;    /* 4*K vertex ids (nvert*K, see nvert); local vertex order fixes node/face order */
    const int8_t* face_bc;  /* 4*K DGTD_BC_* codes (faces*K, see nvert; 1-3 boundary, 4-7 interior faces),
                             * NULL = none */
    const double* eps_r;    /* K relative permittivity, NULL = 1 */
    const double* mu_r;     /* K relative permeability, NULL = 1 */
    int32_t dim;            /* 0 or 3: tetrahedra.  2: triangles -- the 2D hesthaven
                             * operator (HesthavenEvolution.cpp:450-542 on a 2D mesh,
                             * d/dz = 0 and n_z = 0, DGOperatorFactory.h:428-437):
                             * EToV and face_bc hold 3 entries per element, edge f
                             * joins local vertices (f, f+1 mod 3) as in MFEM's
                             * triangle, codes 1-7, vx keeps 3 coordinates per
                             * vertex with z ignored; the state keeps all six
                             * components, Np = (p+1)(p+2)/2. */
    const double* sigma;    /* K bulk conductivity, normalised as the reference's
                             * Material (sigma_SI * Z0, Material.h:30, driver.cpp:962-963),
                             * NULL = 0.  Only with DGTD_KERNEL_CSR (and on hexahedra): the Global operator's
                             * -M_eps^-1 M_sigma E term (collectGlobalConductiveOperator,
                             * DGOperatorFactory.h:1351-1361), -sigma/eps_r E per element. */
    int32_t nvert;          /* vertices per element.  With dim = 2: 0 or 3 triangles, 4
                             * quadrilaterals -- parallelograms; EToV and face_bc hold 4
                             * entries per element, edge f joins local vertices (f, f+1
                             * mod 4), nodes are the (p+1)^2 Gauss-Lobatto tensor points
                             * in MFEM's lexicographic order.  With dim = 0 or 3: 0 or 4
                             * tetrahedra, 8 hexahedra -- parallelepipeds, vertices in
                             * MFEM's order (0 (0,0,0), 1 (1,0,0), 2 (1,1,0), 3 (0,1,0),
                             * 4..7 the same at z = 1 of the unit cube); face_bc holds 6
                             * entries per element for MFEM's faces (vertex cycles
                             * (3,2,1,0), (0,1,5,4), (1,2,6,5), (2,3,7,6), (3,0,4,7),
                             * (4,5,6,7)), nodes are the (p+1)^3 Gauss-Lobatto tensor
                             * points, node i + (p+1) j + (p+1)^2 k at (x_i, x_j, x_k).
                             * Both run the Global operator's flux, the reference's default
                             * operator on its quadrilateral and hexahedral cases
                             * (ExtensiveCasesTest.cpp:1003,1082, Solver3DTest.cpp:57-215):
                             * quadrilaterals assembled (DGTD_KERNEL_CSR; AUTO selects it),
                             * hexahedra assembled (DGTD_KERNEL_CSR) or matrix-free
                             * (DGTD_KERNEL_VALU, which AUTO selects; sigma allowed). */
} dgtd_mesh_desc;

/* Options; mirrors EvolutionOptions (src/evolution/EvolutionOptions.h:13-20)
 * plus the implementation's precision and integrator. */
typedef struct {
    double alpha;       /* upwind weight: 1 upwind, 0 centred (EvolutionOptions::alpha) */
    int32_t precision;  /* 64 or 32 */
    int32_t integrator; /* DGTD_RK4 or DGTD_LSERK4 */
    int32_t device;     /* HIP device ordinal */
    int32_t kernel;     /* DGTD_KERNEL_*: stage-kernel variant, 0 = autotune */
} dgtd_opts;

/* Stage-kernel variants (both fuse gather, flux, curl, LIFT and the RK update). */
#define DGTD_KERNEL_AUTO 0 /* time the available variants on this mesh at create, keep the fastest */
#define DGTD_KERNEL_VALU 1 /* per-node vector FMA, reference matrices in LDS; on hexahedra the
                            * matrix-free tensor-product stage with the Global operator's flux */
#define DGTD_KERNEL_MFMA 2 /* matrix cores (v_mfma 16x16x4), 8-element wave tasks */
/* The assembled operator: the reference's Global operator path
 * (GlobalEvolution::Mult, GlobalEvolution.cpp:628-1088, a CSR SpMV over the
 * matrix DGOperatorFactory::buildGlobalOperator assembles,
 * DGOperatorFactory.h:1428-1600).  The matrix holds the same semi-discrete
 * operator as the matrix-free variants, except that SMA boundary faces weight
 * their upwind terms by 1 whatever alpha, as the Global operator does
 * (DGOperatorFactory.h:485-495, 521-531, 557-567; identical at alpha = 1),
 * plus bulk conductivity; one fused kernel does SpMV and stage update.
 * Chosen explicitly (never by autotune) on tetrahedra, triangles and
 * hexahedra (whose AUTO takes their matrix-free stage, DGTD_KERNEL_VALU);
 * the only kernel of quadrilaterals (AUTO selects it there). */
#define DGTD_KERNEL_CSR 3

const char* dgtd_last_error(void);
const char* dgtd_version(void);

/* Build the operator: geometry, connectivity, reference matrices (FP64 setup),
 * device upload.  Replaces HesthavenEvolution::HesthavenEvolution
 * (HesthavenEvolution.cpp:315-437) and Connectivities (HesthavenEvolutionMethods.cpp:501-534). */
int dgtd_op_create(const dgtd_mesh_desc* mesh, const dgtd_opts* opts, dgtd_op** out);
void dgtd_op_destroy(dgtd_op* op);

/* Sizes: ndofs = 6*Np*K, the operator height (HesthavenEvolution.cpp:316). */
int dgtd_op_ndofs(const dgtd_op* op, int64_t* ndofs);
int dgtd_op_np(const dgtd_op* op, int32_t* Np);

/* Largest fscale over all element faces; Solver::estimateTimeStep reads it
 * through getHesthavenElement(e).fscale (Solver.cpp:337-340). */
int dgtd_op_max_fscale(const dgtd_op* op, double* out);

/* The reference's automatic time step (estimateTimeStep, Solver.cpp:302-353)
 * for this operator's mesh under the named evolution operator
 * (DGTD_OPERATOR_*): on tetrahedra under hesthaven cfl / (max fscale * p^2);
 * otherwise cfl * 0.75 * (2/3) * rmin * min_e vol_e / (surface_e / 2)
 * (calcMeshTimeStep and getTimeStepScale, Solver.cpp:207-300; rmin is the
 * first Gauss-Lobatto spacing of order p on [-1, 1], speed of light 1),
 * halved on quadrilaterals, divided by 6 on hexahedra and by 0.8 on
 * tetrahedra under Global. */
int dgtd_op_estimate_time_step(const dgtd_op* op, int32_t evolution_operator, double cfl, double* dt);
/* The same from a mesh description, on the host (no device work). */
int dgtd_mesh_estimate_time_step(const dgtd_mesh_desc* mesh, int32_t evolution_operator, double cfl, double* dt);

/* Physical node positions, 3*Np*K doubles: x block, y block, z block. */
int dgtd_op_node_positions(const dgtd_op* op, double* xyz);

/* out = L(in) -- HesthavenEvolution::Mult (HesthavenEvolution.cpp:450-542).
 * Host pointers, 6*Np*K doubles each; `t` mirrors TimeDependentOperator::SetTime
 * (used by the reference only for TF/SF sources: the plane waves of
 * dgtd_op_set_planewaves are evaluated at t). */
int dgtd_op_mult(dgtd_op* op, double t, const double* in, double* out);

/* Same on device pointers in the operator's precision, on `stream` (a
 * hipStream_t; NULL = the handle's stream).  Asynchronous. */
int dgtd_op_mult_device(dgtd_op* op, double t, const void* d_in, void* d_out, void* stream);

/* Device-resident state: upload / download the state vector (host FP64). */
int dgtd_op_set_state(dgtd_op* op, const double* x, double t);
int dgtd_op_get_state(dgtd_op* op, double* x, double* t);

/* Advance the resident state by nsteps steps of size dt with the handle's
 * integrator; stage updates are fused into the operator kernel (one read and
 * one write of each state buffer per stage).  Mirrors Solver::step
 * (Solver.cpp:535-551) without the final-time truncation. */
int dgtd_op_advance(dgtd_op* op, double dt, int32_t nsteps);

/* Solver::run (Solver.cpp:483-533) on the resident state: steps of dt while
 * t <= final_time - 1e-8*dt, the last truncated to final_time - t (:537).
 * Writes the number of steps taken. */
int dgtd_op_run(dgtd_op* op, double dt, double final_time, int32_t* steps);

/* Time of the resident state (TimeDependentOperator::GetTime after a run). */
int dgtd_op_time(const dgtd_op* op, double* t);

/* Host convenience: x (host FP64) advanced by nsteps steps; t updated. */
int dgtd_op_step(dgtd_op* op, double* x, double* t, double dt, int32_t nsteps);

/* *flag = 1 if a time step since the last dgtd_op_set_state stored a
 * non-finite state entry or one above 1e20 in magnitude, else 0: the per-step
 * stability test of Solver::run (Solver.cpp:499-516; the reference tests the
 * state norm, non-finite or above 1e20, and prints a warning while the run
 * goes on, as dgtd_op_run does).  Tested in the fused last stage of every
 * step, at no extra pass over the state. */
int dgtd_op_unstable(dgtd_op* op, int32_t* flag);

/* Allocation-free variant of dgtd_op_norms.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The same three FP64 values of a device state in the operator's precision
 * (d_state; NULL = the resident state) are written to d_out3 (3 device
 * doubles), asynchronously on `stream` (NULL = the handle's stream).  The
 * reduction's partial sums live in a static device array of the library
 * (3 x 1024 doubles per GPU, resolved at dgtd_op_create); an operator whose
 * reduction needs more than 1024 partial blocks is refused with DGTD_E_ARG
 * (none does: dgtd_op_create caps the count at 1024).  No hipMalloc, hipFree
 * or allocating library call is made.  Calls that overlap on one GPU share
 * that array and race: serialise them (one stream, or events). */
int dgtd_op_norms_noalloc(dgtd_op* op, const void* d_state, double* d_out3, void* stream);

/* FP64 reductions on the resident state (fixed-order, no atomics):
 * out[0] = DOF 2-norm (Fields::getNorml2, Fields.h:21),
 * out[1] = energy 0.5*sum_e J_e (eps E^T M E + mu H^T M H),
 * out[2] = 1 if any entry is non-finite, else 0. */
int dgtd_op_norms(dgtd_op* op, double* out3);

/* Plane waves for the TF/SF faces (Planewave with a Gaussian or modulated
 * Gaussian magnitude, Function.h:329-418 and :45-140; built by
 * buildGaussianPlanewave / buildModulatedGaussianPlanewave, driver.cpp:483-520):
 *   u = k.x - t (c = 1),  g = exp(-(u - mean)^2 / (2 spread^2))
 *   [* cos(2 pi frequency (u - mean)) for DGTD_PLANEWAVE_MODULATED],
 *   E_inc = g pol,  H_inc = g (k x pol),  pol and k normalised here.
 * mean is the projected 1D mean (k.mean of the case's 3D mean); frequency is
 * normalised (f / c).  The incident field is evaluated at each stage time
 * (LSERK t + c_s dt; RK4 t, t + dt/2, t + dt/2, t + dt; Mult: its t).
 * Replaces the previous set; n = 0 switches the injection off. */
#define DGTD_PLANEWAVE_GAUSSIAN 0
#define DGTD_PLANEWAVE_MODULATED 1
typedef struct {
    int32_t kind;
    double polarization[3];
    double propagation[3];
    double spread;
    double mean;
    double frequency;
} dgtd_planewave;
int dgtd_op_set_planewaves(dgtd_op* op, const dgtd_planewave* waves, int32_t n);

/* Point probes (PointProbe, src/components/Probes.h:132-159, evaluated by
 * ProbesManager::updateProbe, src/solver/ProbesManager.cpp:375-406): the
 * six field components of the nodal solution at n points.  probe_set
 * locates each point (the lowest-index element whose closure holds it; the
 * DG field may jump across element faces) and stores its interpolation
 * weights on the device; elem[p] (may be NULL) receives the element or -1
 * for a point outside the mesh.  A new call replaces the previous set. */
int dgtd_op_probe_set(dgtd_op* op, int32_t n, const double* xyz, int32_t* elem);
/* out[6p + c] = component c (Ex,Ey,Ez,Hx,Hy,Hz) of the resident state at
 * probe p, FP64 (NaN for a probe outside the mesh).  Synchronous. */
int dgtd_op_probe_sample(dgtd_op* op, double* out);

/* Kernel timing of the last dgtd_op_advance/dgtd_op_run call: average
 * operator-kernel duration in microseconds over its launches, measured with
 * HIP events on the handle's stream (0 when timing is off). */
int dgtd_op_set_timing(dgtd_op* op, int32_t on);

/* on != 0: time steps of dgtd_op_advance / dgtd_op_run / dgtd_op_step are
 * replayed from a HIP graph of one step (captured per dt on first use, not
 * used while plane waves are set).  Default 0: every stage kernel is launched
 * directly.  Results are identical either way. */
int dgtd_op_set_graphs(dgtd_op* op, int32_t on);
int dgtd_op_last_kernel_us(const dgtd_op* op, double* avg_us, int32_t* launches);

/* Stage-kernel variant in use (DGTD_KERNEL_VALU, _MFMA or _CSR). */
int dgtd_op_kernel_variant(const dgtd_op* op, int32_t* variant);

/* The stage kernel the autotuner chose (no counterpart in the reference).
 * With DGTD_KERNEL_AUTO every candidate is timed at create on the zeroed
 * state (the fused LSERK stage for LSERK4 operators, Mult otherwise; two
 * rounds, three launches then about 1 ms of launches each, the lower mean
 * kept) and the fastest kept; the choice is
 * cached for the process per (element shape, order, precision, stage mode,
 * floor(log2 K)), and later operators of that kind take it without timing.
 * Candidates: 0 the per-node stage; 1 + 2c + o the matrix-core stage in
 * configuration c (0: up to 12 FP32 / 8 FP64 waves per block, exterior
 * traces gathered 2 k-steps ahead, residual read at the update; 1: up to 8 /
 * 4 waves with twice the registers per wave, gathered 16 FP32 / 4 FP64
 * k-steps ahead, the first residual pass read during the flux loop; FP32
 * stages with the per-lane tail of output nodes, orders 3 and 4, run at most
 * 8 waves in either configuration) with its 8-element tasks in Morton order
 * (o = 0) or element order (o = 1). */
typedef struct {
    int32_t variant;          /* DGTD_KERNEL_VALU, _MFMA or _CSR */
    int32_t config;           /* matrix-core configuration */
    int32_t waves_per_block;
    int32_t gather_depth;
    int32_t early_residual;
    int32_t task_order;       /* 1 Morton, 0 element order */
    int32_t from_cache;       /* 1: taken from the process cache, not timed */
    int32_t ncandidates;
    double candidate_us[8];   /* per candidate, microseconds per launch (0: not timed or does not fit) */
} dgtd_tuning;
int dgtd_op_tuning(const dgtd_op* op, dgtd_tuning* out);

/* Pins the stage kernel to one candidate of the autotuner (no counterpart in
 * the reference): the candidate numbered as above (0 .. ncandidates-1).
 * DGTD_E_ARG when it does not fit this operator (no matrix-core variant at
 * this order and precision, element order without a Morton permutation) or
 * the operator is assembled (DGTD_KERNEL_CSR).  The matrix-core candidates
 * compute the same arithmetic in the same order, so they store identical
 * results; the per-node candidate sums in another order and agrees to
 * rounding (FP64 to 1e-12 of the oracle).  Callers pin one for reproducible
 * timing or bits. */
int dgtd_op_set_candidate(dgtd_op* op, int32_t candidate);

/* Algorithmic HBM bytes and flops per element per stage for this operator. */
int dgtd_op_work_per_element(const dgtd_op* op, double* bytes, double* flops);

/* Structured cube generator: n^3 hexahedra on [0,L]^3, each split into six
 * tetrahedra around the main diagonal; negative tetrahedra get local vertices
 * 2 and 3 swapped.  side_bc[6] = codes for x=0, x=L, y=0, y=L, z=0, z=L.
 * Sizes: vx 3*(n+1)^3, EToV 4*6n^3, face_bc 4*6n^3 (any may be NULL to query). */
int dgtd_cube_mesh(int32_t n, double L, const int8_t* side_bc, double* vx, int32_t* EToV,
                   int8_t* face_bc);

/* Initial fields, evaluated at the nodes (the reference's L2 nodal
 * projection of InitialField sources, SourcesManager.cpp:16-36 with
 * InitialField::eval, Sources.cpp:37-57):
 *   value(x) = f(x - center) * polarization[d] added to component (field, d)
 *   gaussian: f = exp(-sum_{k<dimension} y_k^2 / (2 spread^2))  (Function.h:45-100)
 *   resonant: f = prod_{k<nmodes} sin(modes[k] pi y_k)          (Function.h:145-168)
 * Several sources add up, as setInitialFields does. */
#define DGTD_FIELD_GAUSSIAN 0
#define DGTD_FIELD_RESONANT 1
/* besselj6_2D: f = J_6(a6 r) cos(6 phi), r and phi of (y_0, y_1)        (Function.h:170-194)
 * besselj6_3D: f = j_6(a6 r) Y_6^6(theta, phi) real part, with GSL's
 *              normalised associated Legendre function               (Function.h:196-222)
 * a6 = 13.589290170541217 (first zero of J_6); both at t = 0. */
#define DGTD_FIELD_BESSELJ6_2D 2
#define DGTD_FIELD_BESSELJ6_3D 3
typedef struct {
    int32_t kind;            /* DGTD_FIELD_* */
    int32_t field;           /* 0 electric, 1 magnetic */
    double center[3];
    double polarization[3];
    double spread;           /* gaussian */
    int32_t dimension;       /* gaussian: 1..3 coordinates used */
    int32_t nmodes;          /* resonant: 1..3 */
    int32_t modes[3];
} dgtd_initial_field;

/* x (6*Np*K doubles, reference layout) = sum of the sources at the nodes. */
int dgtd_project_initial_fields(const dgtd_mesh_desc* mesh, const dgtd_initial_field* sources,
                                int32_t nsources, double* x);

/* ------------------------------------------------------------------------
 * Case files: the reference's JSON launcher input (driver.cpp:531-706,
 * 940-1035, 1242-1300) as far as the hesthaven and assembled operators run
 * it.
 *   solver_options: upwind_alpha, time_step (0 = the reference's automatic rule,
 *                   dgtd_case_time_step), final_time,
 *                   cfl, order, evolution_operator, basis_type (GaussLobatto),
 *                   ode_type (RK4)
 *   model: filename (Gmsh 2.2 ASCII .msh or MFEM v1.0 .mesh, relative to the
 *          case file; driver.cpp:1083-1188; tetrahedra or hexahedra, triangles
 *          or quadrilaterals), refinement (uniform, 1:8 in 3D),
 *          materials (relative_permittivity / _permeability / bulk_conductivity
 *          per element tag; a conductive case needs DGTD_KERNEL_CSR),
 *          boundaries (PEC / PMC / SMA per tag; a tag on interior faces
 *          becomes the interior condition DGTD_BC_INT_*)
 *   sources: "initial" gaussian / resonant / besselj6_2D / besselj6_3D fields; "planewave" TF/SF
 *            sources (Gaussian or modulated Gaussian, tags = TF/SF surface)
 *   probes: point positions, exporter steps, domain snapshots
 * Dipole sources, SGBC and non-RK4 integrators are refused
 * with DGTD_E_UNSUPPORTED.
 * ------------------------------------------------------------------------ */
typedef struct dgtd_case dgtd_case;

#define DGTD_OPERATOR_MAXWELL 0   /* EvolutionOperatorType (EvolutionOptions.h:7-11) */
#define DGTD_OPERATOR_GLOBAL 1
#define DGTD_OPERATOR_HESTHAVEN 2

typedef struct {
    int32_t K, nv, order;
    double alpha, time_step, final_time, cfl;
    int32_t evolution_operator;   /* as named in the case (DGTD_OPERATOR_*; "global" when absent):
                                   * the matrix-free hesthaven operator, or the assembled one
                                   * for conductivity, SMA faces with alpha != 1 under Global
                                   * or Maxwell, quadrilaterals and hexahedra */
    int32_t nsources, nprobes;
    int32_t exporter_steps;       /* probes.exporter.steps, 0 if absent */
    int32_t refinement;           /* uniform refinement levels applied */
    int32_t snapshot_steps;       /* probes.domain_snapshot: "steps" (0 if absent), -1 = no snapshots */
    int32_t snapshot_saves;       /* probes.domain_snapshot: "saves", 0 if absent */
    int32_t nplanewaves;          /* "planewave" sources (their tags are TF/SF faces) */
    int32_t dim;                  /* 3: tetrahedral or hexahedral mesh; 2: triangle or
                                   * quadrilateral mesh (dgtd_case_mesh sets nvert = 4 for
                                   * quadrilaterals, 8 for hexahedra, which run on the
                                   * assembled operator; such a case naming the hesthaven
                                   * operator is refused) */
    int32_t nfield_probes;        /* probes.field entries */
} dgtd_case_info;

/* Parse the case file, read and refine its mesh, map tags to boundary codes
 * and materials (buildModel, driver.cpp:1242-1300). */
int dgtd_case_load(const char* json_path, dgtd_case** out);
void dgtd_case_destroy(dgtd_case* c);
int dgtd_case_get_info(const dgtd_case* c, dgtd_case_info* info);
/* Mesh description for dgtd_op_create; its arrays belong to the case. */
int dgtd_case_mesh(const dgtd_case* c, dgtd_mesh_desc* mesh);
/* info.nplanewaves plane waves for dgtd_op_set_planewaves (driver.cpp:565-609;
 * without a "mean" the pulse starts 5*sqrt(2) spreads before the upstream
 * point of the TF/SF surface, driver.cpp:575-588). */
int dgtd_case_planewaves(const dgtd_case* c, dgtd_planewave* out);
/* info.nsources initial fields (buildSources, driver.cpp:531-564). */
int dgtd_case_sources(const dgtd_case* c, dgtd_initial_field* out);
/* Domain snapshots (DomainSnapshotProbe, ProbesManager.cpp:457-495): the
 * case's (refined) mesh in MFEM v1.0 format, and the six field components of
 * a state as MFEM GridFunction files Ex.gf ... Hz.gf (L2 GaussLobatto space of
 * the given order, 16 significant digits) plus time.txt in directory dir. */
int dgtd_case_save_mesh(const dgtd_case* c, const char* path);
/* The case's time step: solver_options.time_step when > 0, else
 * dgtd_mesh_estimate_time_step on the case's mesh under its evolution
 * operator and cfl (Solver.cpp:110-115). */
int dgtd_case_time_step(const dgtd_case* c, double* dt);
int dgtd_save_fields(const char* dir, int32_t order, int64_t ndofs, const double* x, double t);
/* The same with the case's dimension and order (L2_T1_2D_P<p> on triangles). */
int dgtd_case_save_fields(const dgtd_case* c, const char* dir, int64_t ndofs, const double* x, double t);

/* Point probes (buildProbes, driver.cpp:707-864): 3 * info.nprobes
 * coordinates, and per probe its "steps" and "saves" entries (0 if absent;
 * either array may be NULL).  The sampling interval is steps, else
 * max(1, ceil(final_time/dt)/saves), else 1 (driver.cpp:717-733). */
int dgtd_case_probes(const dgtd_case* c, double* xyz, int32_t* steps, int32_t* saves);

/* Field probes (FieldProbe, Probes.h:75-130; driver.cpp:764-778): one field
 * component at a point, written as FieldProbes/FieldProbe<ID>.dat
 * (ProbesManager.cpp:222-259, 344-371).  3 * info.nfield_probes coordinates,
 * "steps" and "saves" as for point probes, and the component 3*field +
 * direction (0..5 = Ex, Ey, Ez, Hx, Hy, Hz); any array may be NULL. */
int dgtd_case_field_probes(const dgtd_case* c, double* xyz, int32_t* steps, int32_t* saves, int32_t* component);

/* ------------------------------------------------------------------------
 * Batched instances: many independent meshes of uneven size and order
 * advance together (no counterpart in the reference, which runs one mesh per
 * process).  Each time step is one persistent launch per LSERK4 stage over a
 * device work list of (instance, element block) items plus a bookkeeping
 * launch, captured once in a hipGraph; instances keep their fields resident
 * in HBM, carry their own dt and final time, and retire on the device with
 * the Solver::run loop rule (Solver.cpp:497,537).
 * ------------------------------------------------------------------------ */
typedef struct dgtd_batch dgtd_batch;

typedef struct {
    double alpha;          /* upwind weight for all instances */
    int32_t precision;     /* 64 or 32 */
    int32_t device;        /* HIP device ordinal */
    int32_t max_instances; /* slots, <= 4096 */
    int32_t reserved;
    int64_t hbm_budget;    /* bytes of instance storage admitted at once (0 = 80% of free HBM) */
} dgtd_batch_opts;

/* Instance states reported by dgtd_batch_info.  FAILED: a step stored a
 * non-finite entry or one above 1e20 in magnitude -- the per-step stability
 * test of Solver::run (Solver.cpp:499-516), which the reference only reports;
 * here the instance stops at that step and keeps its fields for inspection.
 * CANCELLED: stopped by dgtd_batch_cancel. */
#define DGTD_INST_FREE 0
#define DGTD_INST_RUNNING 1
#define DGTD_INST_DONE 2
#define DGTD_INST_FAILED 3
#define DGTD_INST_CANCELLED 4

int dgtd_batch_create(const dgtd_batch_opts* opts, dgtd_batch** out);
void dgtd_batch_destroy(dgtd_batch* b);

/* Device bytes one instance of this mesh holds (3 state buffers, geometry,
 * connectivity) -- the admission measure. */
int dgtd_batch_footprint(const dgtd_mesh_desc* mesh, int32_t precision, int64_t* bytes);

/* Admit a mesh with initial fields x0 (host FP64, reference layout; NULL =
 * zero), time step dt (<= 0: the Hesthaven rule 1/(max fscale p^2)) and final
 * time.  Returns DGTD_E_CAPACITY when the HBM budget or the slots are full.
 * Tetrahedral (dim 0 or 3) and triangle (dim 2) meshes may share a batch:
 * triangles of every order and tetrahedra of orders 2-5 run on the
 * matrix-core stage, order-1 tetrahedra on the per-node stage.  Conductivity
 * and state buffers of 4 GiB or more are refused (DGTD_E_UNSUPPORTED), and so
 * are quadrilateral and hexahedral meshes (nvert 4 with dim 2, nvert 8: the
 * assembled operator only); TF/SF
 * faces (code 7) inject nothing without plane waves (dgtd_batch_add_tfsf). */
int dgtd_batch_add(dgtd_batch* b, const dgtd_mesh_desc* mesh, const double* x0, double dt,
                   double final_time, int32_t* id);

/* dgtd_batch_add with plane waves injected on the mesh's TF/SF faces
 * (HesthavenEvolution::evaluateTFSF, HesthavenEvolution.cpp:97-124; the same
 * waves as dgtd_op_set_planewaves takes), at every stage of the instance's own
 * LSERK4 step times. */
int dgtd_batch_add_tfsf(dgtd_batch* b, const dgtd_mesh_desc* mesh, const double* x0, double dt,
                        double final_time, const dgtd_planewave* waves, int32_t nwaves, int32_t* id);

/* Launch nsteps time steps of all running instances (asynchronous). */
int dgtd_batch_step(dgtd_batch* b, int32_t nsteps);

/* Wait for the device, then report instances that ended since the last poll
 * (done, failed or cancelled; up to cap ids) and the number still running. */
int dgtd_batch_poll(dgtd_batch* b, int32_t* done_ids, int32_t cap, int32_t* ndone, int32_t* nrunning);

/* Per instance: time, steps taken, DGTD_INST_* state, dt in use. */
int dgtd_batch_info(dgtd_batch* b, int32_t id, double* t, int32_t* steps, int32_t* state, double* dt);

/* Fields of an instance (host FP64, reference layout). */
int dgtd_batch_get_state(dgtd_batch* b, int32_t id, double* x);

/* Stop a running instance before its final time (no counterpart in the
 * reference, which runs one case per process): it leaves the work list at the
 * next step, keeps its fields as of the last completed step and is reported
 * once by dgtd_batch_poll as DGTD_INST_CANCELLED.  An instance that already
 * ended is left as it is. */
int dgtd_batch_cancel(dgtd_batch* b, int32_t id);

/* Release an instance's device storage and slot. */
int dgtd_batch_remove(dgtd_batch* b, int32_t id);

/* Bytes admitted, running instances, work-list items, resident blocks of the
 * persistent stage launch. */
int dgtd_batch_stats(dgtd_batch* b, int64_t* bytes_in_use, int32_t* running, int32_t* items,
                     int32_t* grid);

/* Host-only introspection of the setup (no GPU needed). */

/* Reference matrices of order p on the Hesthaven tetrahedron (volume 4/3),
 * nodes in the MFEM L2 GaussLobatto order (HesthavenEvolution.cpp:46-81,365):
 * rst 3*Np, M/Dr/Ds/Dt Np*Np, lift Np*4Nfp (row-major), fmask 4*Nfp; any NULL. */
int dgtd_ref_element(int32_t order, double* rst, double* M, double* Dr, double* Ds, double* Dt,
                     double* lift, int32_t* fmask);

/* Expands the compact connectivity into the reference's (vmapM, vmapP) lists,
 * 4*Nfp*K each, in Connectivities::global order (HesthavenEvolutionMethods.cpp
 * :721-791): element, local face, face node ascending; P = M on boundary faces
 * and on faces with an interior condition (their jump uses the own trace only).
 * face_codes (4*K, may be NULL) receives the code of each face (0 for plain
 * interior faces, 4-6 on both sides of a face with an interior condition).
 * Tetrahedral meshes only (DGTD_E_UNSUPPORTED otherwise). */
int dgtd_mesh_vmaps(const dgtd_mesh_desc* mesh, double alpha, int32_t* vmapM, int32_t* vmapP,
                    int8_t* face_codes);

/* Per-element geometric factors, 27 doubles per element: dr/dx..dt/dz (9),
 * per face nx ny nz fscale/2 (16), 1/eps_r, 1/mu_r.  Tetrahedral meshes only
 * (DGTD_E_UNSUPPORTED otherwise). */
int dgtd_mesh_geometry(const dgtd_mesh_desc* mesh, double alpha, double* geo);

/* The assembled operator DGTD_KERNEL_CSR uploads, on the host in FP64 (the
 * Global operator's matrix, DGOperatorFactory::buildGlobalOperator,
 * DGOperatorFactory.h:1428-1600): ndofs x ndofs CSR, rows and columns in the
 * state layout, mesh->sigma included.  Call with rowptr, col and val NULL to
 * get *nnz, then with arrays of ndofs+1, *nnz and *nnz entries. */
int dgtd_global_assemble(const dgtd_mesh_desc* mesh, double alpha, int64_t* nnz, int64_t* rowptr, int32_t* col,
                         double* val);

/* Point location as dgtd_op_probe_set does it: elem[p] (or -1) and Np nodal
 * interpolation weights per point (u(x_p) = sum_n weights[p*Np + n] u_n). */
int dgtd_mesh_locate(const dgtd_mesh_desc* mesh, int32_t n, const double* xyz, int32_t* elem, double* weights);

#ifdef __cplusplus
}
#endif

#endif
