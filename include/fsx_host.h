/* fsx_host.h -- host C++ side of the fractional-step core, C-ABI.
 *
 * Two layers sit on the host, both calling only the device C-ABI in fsx.h:
 *
 *  fsxh_*  mesh input and topology: what the reference's setup chain does
 *          before the time loop (main, BCH:383-521): readInputFile
 *          (BCH:528-725), setupNonCornerNodes (BCH:954-1320), setupLtoGdof
 *          (BCH:1327-1403), determineVelBCnodes (BCH:1410-1580),
 *          findMonitorPoint (BCH:1644-1668), setupSparseM/G (BCH:1675-2159),
 *          the structural pattern of Z (BCH:3322-3479) and a race-free element
 *          colouring (the role of setupMeshColoring, BCH:853-947). The mesh
 *          generators follow the reference's cube and channel generators
 *          (oldFiles/meshGenerators&Converters/HexaMeshGeneratorIn*.m).
 *  fsxd_*  the driver: timeLoop (BCH:2815-3120) over the fsx_* device calls,
 *          in the reference's call order.
 *
 * BCH = fractionalStep/Cpp/blascoCodinaHuerta.cpp.
 */
#ifndef FSX_HOST_H
#define FSX_HOST_H

#include <stdint.h>
#include "fsx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fsxh_mesh fsxh_mesh;
typedef struct fsxh_topo fsxh_topo;
typedef struct fsxd_solver fsxd_solver;

/* ------------------------------------------------------------ mesh input */
/* Lid-driven cavity on the unit cube, N^3 hex27/hex8 elements, optional sinh
 * clustering (0 = uniform), lid velocity (lid_u, 0, 0) on z = 1. Node,
 * element and BC-face order of the reference cube generator. */
fsxh_mesh* fsxh_mesh_cavity(int N, double cluster, double lid_u);
/* 90-degree bent square duct: n x n cross-section elements on [0,1]^2,
 * n_in elements over a straight inlet of length L_in, n_arc over a bend of
 * centreline radius R, n_out over a straight outlet of length L_out.
 * Parabolic inlet u = peak*16 s(1-s) t(1-t); no-slip walls; natural outlet
 * with the zero-pressure node at the outlet centre. permute != 0 shuffles
 * corner-node and element numbering with std::mt19937_64(seed). */
fsxh_mesh* fsxh_mesh_duct(int n, int n_in, int n_arc, int n_out, double L_in, double R, double L_out, double peak,
                          uint64_t seed, int permute);
/* .inp reader/writer in the layout read by BCH:528-725 */
int fsxh_mesh_read_inp(const char* path, fsxh_mesh** out, char* err, int errlen);
int fsxh_mesh_write_inp(const fsxh_mesh* m, const char* path);
void fsxh_mesh_free(fsxh_mesh* m);
/* info: NE, NCN, nVelFaces, nBC, zeroP (0-based), NGP, maxIter, has_profile */
int fsxh_mesh_info(const fsxh_mesh* m, long long* out8);
/* params: dt, t_ini, t_final, tolerance, converge, density, viscosity, mon x,y,z */
int fsxh_mesh_params(const fsxh_mesh* m, double* out9);
int fsxh_mesh_set_params(fsxh_mesh* m, int NGP, double dt, int maxIter, double tolerance, double converge,
                         double density, double viscosity);
/* arrays: "coord" (NCN*3 f64), "conn" (NE*8 i32, 0-based), "faces" (nVelFaces*3
 * i32, 0-based elem/face/bc), "BCstr" (nBC*3 f64) */
int fsxh_mesh_get(const fsxh_mesh* m, const char* name, void* dst);
/* parabolic inlet profile: bc id, peak, lower corner (y0, z0) of the inlet */
int fsxh_mesh_profile(const fsxh_mesh* m, double* out4);

/* ------------------------------------------------------------ topology */
int fsxh_topology(const fsxh_mesh* m, fsxh_topo** out, char* err, int errlen);
void fsxh_topo_free(fsxh_topo* t);
/* sizes: NE, NCN, NN, NNp, nnzM, nnzG, nnzZ, nBCnodes, ncolors, monPoint, zeroP */
int fsxh_topo_info(const fsxh_topo* t, long long* out11);
/* int arrays: "LtoGnode" (NE*27), "Mrow","Mcol","Grow","Gcol","Zrow","Zcol",
 * "BCnodes","BCwhich","color_ptr","color_order" ; f64: "coord" (NN*3),
 * "BCvalues" (nBCnodes*3) */
long long fsxh_topo_len(const fsxh_topo* t, const char* name);
int fsxh_topo_get(const fsxh_topo* t, const char* name, void* dst);

/* ------------------------------------------------------------ driver */
typedef struct fsxd_options {
  int device;
  int NGP;            /* 1, 8 or 27 */
  int maxIter;        /* inner iterations per step (BCH:553) */
  double tolerance;   /* inner convergence (BCH:554) */
  double converge;    /* steady-state criterion (BCH:555) */
  int solver_mode;    /* fsx_solver_mode */
  double pcg_rtol;
  int pcg_maxit;
  int timing;         /* 1: record per-kernel device times */
} fsxd_options;

typedef struct fsxd_step_stats {
  int step;           /* timeN */
  int inner_iters;
  int pcg_iters;      /* summed over the step's solves */
  int steady;
  double time;        /* timeT */
  double max_acc;
  double mon[4];      /* u, v, w, p at the monitor node */
  double wall_s;      /* host wall time of the step (device synchronised) */
  double pcg_relres;  /* of the last solve */
  int pcg_noconv;     /* the step's solves that stopped at the iteration cap (FSX_ENOCONV; their
                         partial answer is used, as the reference does, BCH:4041-4044) */
  int pad_;
  double pcg_relres_max;  /* largest relative residual of the step's solves */
} fsxd_step_stats;

void fsxd_default_options(fsxd_options* o);
/* Builds the topology, creates the device context, uploads, assembles
 * (step0 + Z) and initialises the state. */
int fsxd_create(const fsxh_mesh* m, const fsxd_options* o, fsxd_solver** out, char* err, int errlen);
void fsxd_free(fsxd_solver* s);
/* Runs up to nsteps time steps (fewer if steady state is reached, as the
 * reference does); returns the number of steps done or a negative status. */
int fsxd_run(fsxd_solver* s, int nsteps, fsxd_step_stats* stats);
/* Back to t_ini and the initial fields (fsx_init_state), keeping the
 * assembled matrices; the solver mode may be changed with fsx_set_solver. */
int fsxd_reset(fsxd_solver* s);
fsx_ctx* fsxd_ctx(fsxd_solver* s);
const fsxh_topo* fsxd_topo(fsxd_solver* s);
const char* fsxd_error(fsxd_solver* s);
/* setup wall times (s): topology, upload, assembly (incl. Z) */
int fsxd_setup_times(fsxd_solver* s, double* out3);

/* ------------------------------------------------------------ output and restart
 * The reference's Tecplot layout (createTecplot, BCH:4249-4482): title,
 * variables and zone lines, NN lines "x y z u v w p" in %.11e with the
 * pressure at non-corner nodes interpolated from the element corners (edge,
 * face, element means), then 8 NE one-based sub-brick connectivity lines
 * (each hex27 split into its 8 octants). The restart reader takes the same
 * file back (readRestartFile, BCH:4214-4242). */
int fsxh_write_tecplot(const fsxh_topo* t, const char* title, const double* Un, const double* Pn, const char* path);
int fsxh_read_restart(const fsxh_topo* t, const char* path, double* Un, double* Pn);
/* the driver's current Un, Pn to a Tecplot file */
int fsxd_write_tecplot(fsxd_solver* s, const char* title, const char* path);
/* start from a restart file instead of the initial BCs (BCH:2795-2796):
 * fsxd_reset, then Un and Pn from the file */
int fsxd_read_restart(fsxd_solver* s, const char* path);

/* ------------------------------------------------------------ direct-solve order
 * The fill-reducing order of FSX_SOLVER_DIRECT's sparse factor (the role of
 * the MATLAB prototype's symamd, ML:1880-1881): nested dissection of the n
 * pressure nodes (coordinates xyz, 3 a node) by recursive coordinate
 * bisection with vertex separators from the symmetric pattern (rp, ci) of
 * Z; node sets of at most `leaf` nodes end the recursion. perm (n entries,
 * may be NULL) receives the order, new index -> pressure id, in postorder of
 * the separator tree; stats4 (may be NULL) the entries of L with dense
 * fronts (pivot blocks' lower triangles included), the factor flops (sum of
 * squared column counts), the number of fronts and the tree's depth. The
 * device factor computes this same order. FSX_EINVAL for an inconsistent
 * pattern or a non-finite coordinate. */
int fsxh_nd_order(int n, const int* rp, const int* ci, const double* xyz, int leaf, int* perm, double* stats4);

/* ------------------------------------------------------------ gradient row classes
 * The row classes through which the step kernels read G and G^T at matrix
 * precision 32 (fsx_set_matrix_classes) where the rows compress: G's
 * nrows x ncols CSR pattern (rp, ci, ascending columns) with g3 = (G1, G2,
 * G3) an entry (BCH:3322-3513's G1, G2, G3 on one pattern). A component
 * below prune_rel times its row's largest magnitude counts as 0, the others
 * are rounded to FP32; rows with the same column offsets from their first
 * column and the same values share a class, and G^T's rows are classed by
 * the sequence of G's class entries they hold, and by their column offsets
 * from their first column too where those repeat (a velocity numbering in
 * grid order: the classes then carry the offsets and no row stores its
 * columns). counts5 (may be NULL) receives the numbers of G's classes,
 * their entries, G^T's classes and theirs, and 1 when G^T's classes carry
 * the offsets (all -1 beyond 2^20 classes or 2^28 entries; the device takes
 * the form only where the class entries are at most an eighth of G's); row_cls
 * (nrows) and trow_cls (ncols) the classes; g3_f32 (3 nnz) the values the
 * class form represents, NaN where G^T's class disagrees with G's. The
 * device builds the same classes (with prune_rel 2^-24). FSX_EINVAL for an
 * inconsistent pattern. */
int fsxh_g_classes(int nrows, int ncols, const int* rp, const int* ci, const double* g3, double prune_rel,
                   int* counts5, int* row_cls, int* trow_cls, float* g3_f32);

/* ------------------------------------------------------------ Z on a tensor-product grid
 * The learned solver's LDS-tiled Z products apply Z as Kronecker products
 * of 1D stencils where the pressure grid (nx x ny x nz nodes, x fastest,
 * corner coordinates xc, 3 a node) is a tensor-product box grid:
 *   Z = Kx (x) My (x) Mz + Mx (x) Ky (x) Mz + Mx (x) My (x) Kz,
 * K = D^T m^-1 D / rho^2 and M = B^T m^-1 B per axis, from the 1D quadratic
 * velocity and linear pressure bases with the NGP rule's points per axis
 * (D: derivative of the velocity basis against the pressure basis, B: the
 * bases' product, m: the row-sum lumped 1D velocity mass), i.e. G_x =
 * -(1/rho) D_x (x) B_y (x) B_z of BCH:3205-3207 and the lumped mass of
 * BCH:3262-3278 factored. The form is checked against Z's rows (rp, ci, v
 * of the n = nx ny nz nodes, structural zeros dropped) within 1e-10 of each
 * row's largest magnitude, the pinned node zero_p's diagonal (its penalty)
 * aside. coef ((nx + ny + nz) x 10, may be NULL) receives per axis and
 * grid point K's five coefficients (offsets -2..2) then M's; info3 the
 * largest relative mismatch (-1 when the coordinates do not separate), the
 * pinned diagonal's difference and 1 when the form holds. The device builds
 * the same tables. FSX_EINVAL for inconsistent sizes. */
int fsxh_z_tensor(int nx, int ny, int nz, const double* xc, int ngp, double rho, int n, const int* rp, const int* ci,
                  const double* v, int zero_p, double* coef, double* info3);

#ifdef __cplusplus
}
#endif
#endif /* FSX_HOST_H */
