/*
 * dgtd_oracle.h -- CPU FP64 restatement of OpenSEMBA/dgtd's "hesthaven"
 * DGTD Maxwell operator and the Runge-Kutta drivers around it.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the
 * HIP implementation in dgtd_amd/.  Only tests/, __graft_entry__.smoke() and
 * bench.py's CPU-baseline leg load it; the product never links or calls it.
 *
 * Every routine names the reference file:line it restates.  Parity status of
 * each function (what pins it) is listed in oracle_pins below and in the
 * tests under tests/test_oracle_*.py.
 *
 * State layout (the reference's Fields.h:48-56): x[c*N + e*Np + i] with
 * c in {Ex,Ey,Ez,Hx,Hy,Hz}, N = Np*K.
 */
#ifndef DGTD_ORACLE_H
#define DGTD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_op oracle_op;

/* Reference-element sizes. */
int oracle_np(int order);
int oracle_nfp(int order);

/* Reference element on the Hesthaven tetrahedron (volume 4/3), nodes in the
 * MFEM L2 GaussLobatto order.  Outputs (row-major):
 *   rst   : 3*Np   (r,s,t of each node, r = 2*xi - 1 etc.)
 *   M     : Np*Np  reference mass matrix
 *   Dr,Ds,Dt : Np*Np reference differentiation matrices
 *   emat  : Np*4Nfp  face-mass matrix (Hesthaven convention)
 *   lift  : Np*4Nfp  LIFT = M^-1 * emat
 *   fmask : 4*Nfp    local node index of each face node
 * Any pointer may be NULL. */
void oracle_ref_element(int order, double* rst, double* M, double* Dr,
                        double* Ds, double* Dt, double* emat, double* lift,
                        int* fmask);

/* 1D Gauss-Lobatto-Legendre points on [0,1], p+1 of them. */
void oracle_gll01(int p, double* x);

/* Operator over an affine tetrahedral mesh.
 *   vx      : 3*nv vertex coordinates
 *   EToV    : 4*K vertex ids; local vertex order fixes node and face order
 *   face_bc : 4*K, face f is opposite local vertex f.  On boundary faces
 *             0 (zero jump), 1 PEC, 2 PMC, 3 SMA; on interior faces 0, the
 *             interior conditions 4 PEC, 5 PMC, 6 SMA, or 7 TF/SF surface
 *             (either side; the lower-index element is the scattered-field side).
 *   eps,mu  : K relative permittivity/permeability, NULL means 1
 * Returns NULL on error (message via oracle_last_error). */
/* Plane wave for TF/SF injection (Planewave, Function.h:329-418): kind 0
 * Gaussian, 1 modulated Gaussian; pol and dir are normalised by
 * oracle_set_planewaves; mean is the projected 1D mean k.mean. */
typedef struct {
    int kind;
    double pol[3], dir[3];
    double spread, mean, freq;
} oracle_planewave;

oracle_op* oracle_create(int K, int order, const double* vx, int nv,
                         const int* EToV, const int* face_bc,
                         const double* eps, const double* mu, double alpha);
void oracle_destroy(oracle_op* op);
const char* oracle_last_error(void);

/* Plane waves injected on the TF/SF faces (replaces the previous set). */
int oracle_set_planewaves(oracle_op* op, const oracle_planewave* w, int n);

/* oracle_create with global_sma != 0: SMA faces, boundary (code 3) and
 * interior (code 6), take upwind weight 1 whatever alpha, as the reference's
 * Global operator assembles them (SMA coefficient 1 in bdrCentCoeff /
 * bdrUpwindCoeff, DGOperatorFactory.h:21-35; boundary sub-operators
 * :485-567; interior-boundary sub-operators :575-665, where SMA gets coeff
 * 1.0 and PEC/PMC c*alpha): such a face behaves as a hesthaven face with
 * alpha = 1 (boundary SMA jump -1, flux weight 1), so alpha = 0 with SMA
 * faces is allowed. */
oracle_op* oracle_create_global(int K, int order, const double* vx, int nv,
                                const int* EToV, const int* face_bc,
                                const double* eps, const double* mu, double alpha);

int64_t oracle_ndofs(const oracle_op* op); /* 6*Np*K */
double oracle_max_fscale(const oracle_op* op);

/* Physical node positions, 3*Np*K: x at [e*Np+i], y at [N + ...], z at [2N+...]. */
void oracle_node_positions(const oracle_op* op, double* xyz);

/* (vmapM, vmapP) in the reference's order, 4*Nfp*K each. */
void oracle_connectivity(const oracle_op* op, int* vmapM, int* vmapP);

/* Number of boundary faces of each kind (PEC, PMC, SMA) and their node
 * lists: out_nodes receives, for kind `kind`, nfaces*Nfp DOF ids. */
int oracle_bc_faces(const oracle_op* op, int kind, int* out_nodes);

/* Per-element geometry: vol[K], fscale[4K], normals[12K] (nx,ny,nz per face). */
void oracle_geometry(const oracle_op* op, double* vol, double* fscale,
                     double* normals);

/* out = L(in); the reference's HesthavenEvolution::Mult. */
void oracle_mult(const oracle_op* op, const double* in, double* out);
/* The same at time t (the TF/SF incident field depends on it). */
void oracle_mult_t(const oracle_op* op, const double* in, double* out, double t);

/* One classic RK4 step in mfem::RK4Solver's stage order. */
void oracle_rk4_step(const oracle_op* op, double* x, double t, double dt);

/* One 5-stage low-storage RK4 step (Carpenter & Kennedy 1994). */
void oracle_lserk4_step(const oracle_op* op, double* q, double* res, double t,
                        double dt);

/* Runs the reference's Solver::run loop (Solver.cpp:483-551) with the given
 * integrator (0 = RK4 in MFEM order, 1 = LSERK4); returns steps taken. */
int oracle_run(const oracle_op* op, double* x, double dt, double final_time,
               int integrator, double* t_out);

/* Energy 0.5*sum_e Vol_e/Vref * (eps E^T M E + mu H^T M H). */
double oracle_energy(const oracle_op* op, const double* x);

/* Number of OpenMP threads oracle_mult uses (1 when built without OpenMP). */
int oracle_threads(void);
/* Sets that number for later calls (n > 0). */
void oracle_set_threads(int n);

#ifdef __cplusplus
}
#endif

#endif
