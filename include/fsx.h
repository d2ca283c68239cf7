/* fsx.h -- C-ABI of the MI355X-native fractional-step core (device side).
 *
 * This is the drop-in boundary that replaces the reference's GPU module,
 * fractionalStep/Cpp/CUDAcodes.cu (CU), whose C++-linkage functions the
 * reference driver fractionalStep/Cpp/blascoCodinaHuerta.cpp (BCH) calls
 * (prototypes BCH:365-377). Each entry point below names the reference
 * function it replaces. Data crosses as plain pointers and sizes; the device
 * state lives in an opaque context instead of ~60 globals (BCH:141-322,
 * CU:47-97). Every call returns FSX_OK (0) or an error code; the message is
 * available from fsx_last_error(). Nothing blocks on stdin or exits.
 *
 * Conventions (the reference's layouts, BCH:1327-1403, 1807-1811):
 *   velocity vectors are [u(0..NN-1) | v | w], FP64, in the reference's node
 *   numbering (the device keeps its own node order and layout and converts
 *   at every call that takes or returns velocity-sized data or M/G values);
 *   pressure is indexed by corner node (NNp = NCN), FP64;
 *   CSR row pointers and columns are int32, 0-based, columns sorted.
 * One context per host thread; all work is queued on the context's stream.
 */
#ifndef FSX_H
#define FSX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fsx_ctx fsx_ctx;

enum fsx_status {
  FSX_OK = 0,
  FSX_EINVAL = 1,   /* bad argument or size */
  FSX_EHIP = 2,     /* HIP runtime error */
  FSX_ESTATE = 3,   /* call out of order (e.g. step before assembly) */
  FSX_ENOCONV = 4,  /* linear solver hit maxit before rtol */
  FSX_ENOMEM = 5,   /* device allocation failed */
  FSX_ENUMERIC = 6  /* NaN/Inf met in a solver */
};

/* CSR kinds for fsx_upload_csr */
enum fsx_csr_kind {
  FSX_CSR_M = 0,   /* NN x NN pattern shared by M, K, A  (setupSparseM BCH:1675-1921) */
  FSX_CSR_G = 1,   /* NN x NNp pattern shared by G1..G3  (setupSparseG BCH:1928-2159) */
  FSX_CSR_Z = 2    /* NNp x NNp structural pattern of G^T G (calculateZ BCH:3322-3479) */
};

/* Vector / array ids for fsx_download and fsx_upload_vec. Sizes in doubles. */
enum fsx_vec_id {
  FSX_VEC_UN = 0,            /* 3NN */
  FSX_VEC_UNP1 = 1,          /* 3NN */
  FSX_VEC_UNP1_PREV = 2,     /* 3NN */
  FSX_VEC_UNPHALF = 3,       /* 3NN */
  FSX_VEC_UNPHALF_PREV = 4,  /* 3NN */
  FSX_VEC_ACC = 5,           /* 3NN */
  FSX_VEC_KACC = 6,          /* 3NN */
  FSX_VEC_R1 = 7,            /* 3NN (after fsx_convection: -A*UnpHalf_prev; after fsx_step1: full R1) */
  FSX_VEC_PN = 8,            /* NNp */
  FSX_VEC_PNP1 = 9,          /* NNp */
  FSX_VEC_PNP1_PREV = 10,    /* NNp */
  FSX_VEC_PDOT = 11,         /* NNp */
  FSX_VEC_R2 = 12,           /* NNp */
  FSX_VEC_MDINV = 13,        /* 3NN (replicated x3 like BCH:3268-3270) */
  FSX_VEC_MDORIGINV = 14,    /* 3NN */
  FSX_VAL_K = 15,            /* nnz(M) values of K */
  FSX_VAL_M = 16,            /* nnz(M) values of the consistent mass M */
  FSX_VAL_G1 = 17,           /* nnz(G) */
  FSX_VAL_G2 = 18,
  FSX_VAL_G3 = 19,
  FSX_VAL_Z = 20,            /* nnz(Z), penalty applied */
  FSX_VEC_COUNT = 21
};

/* Pressure-solver modes for fsx_set_solver (CUSP_CG_solver CU:418-451 and
 * MKL_CG_solver BCH:3981-4055 are the FP64 Jacobi form). */
enum fsx_solver_mode {
  FSX_SOLVER_FP64_JACOBI = 0,  /* FP64 Jacobi-PCG, x0 = 0, ||r|| <= rtol ||b|| */
  FSX_SOLVER_MIXED_IR = 1,     /* FP32 inner Jacobi-PCG, FP64 residual + refinement, same stop test */
  FSX_SOLVER_LEARNED = 2,      /* flexible PCG with the learned preconditioner, Jacobi-PCG fallback */
  FSX_SOLVER_DIRECT = 3        /* Cholesky of Z factored once, two triangular solves per step (the MATLAB
                                  prototype's pressure step, ML:1876-1885, 2021-2030); rtol bounds the
                                  FP64 residual it must meet */
};

/* Initial guess of the pressure solve for fsx_set_solver_x0 */
enum fsx_x0_policy {
  FSX_X0_ZERO = 0,      /* x0 = 0, the CPU path (MKL RCI CG, BCH:4002-4004); the default */
  FSX_X0_PREVIOUS = 1,  /* the previous solution, the GPU path (CUSP keeps Pdot, CU:268, 425-442; Q6) */
  FSX_X0_PROJECTION = 2 /* the Z-norm projection onto the last 16 solutions (no reference counterpart) */
};

/* Velocity mass for fsx_set_mass */
enum fsx_mass_mode {
  FSX_MASS_LUMPED = 0,      /* row-sum lumped M, the reference's update (BCH:3262-3295, 3781-3783, 3960-3962) */
  FSX_MASS_CONSISTENT = 1,  /* consistent M, one 3-RHS Jacobi-PCG per update (no reference counterpart),
                               M applied element by element */
  FSX_MASS_CONSISTENT_CSR = 2   /* the same solves with M applied from its assembled CSR values */
};

/* ---------------------------------------------------------------- lifecycle */
/* replaces selectCUDAdevice (CU:126-154) and the cuSPARSE/cuBLAS setup (CU:171-178) */
int fsx_create(int device, fsx_ctx** out);
void fsx_destroy(fsx_ctx* ctx);
const char* fsx_last_error(const fsx_ctx* ctx);
/* device memory currently held by the context, in bytes */
size_t fsx_device_bytes(const fsx_ctx* ctx);

/* ---------------------------------------------------------------- uploads
 * Together these replace initializeAndAllocateGPU (CU:161-274). Host arrays
 * are caller-owned and copied. */
/* LtoGnode: NE*27 global node ids per element in the reference's local order
 * (BCH:2312-2343); corner_coord: NNp*3 corner coordinates (BCH:582-585).
 * color_order: NE element ids grouped by colour, color_ptr: ncolors+1 offsets;
 * no two elements of one colour may share a node (checked on the device side
 * inputs by the host library; see fsxh_color_elements). */
int fsx_upload_mesh(fsx_ctx* ctx, int NE, int NN, int NNp, const int* LtoGnode, const double* corner_coord,
                    int ncolors, const int* color_ptr, const int* color_order);
/* rowptr has nrows+1 entries, col rowptr[nrows] entries. */
int fsx_upload_csr(fsx_ctx* ctx, int kind, int nrows, const int* rowptr, const int* col);
/* velocity BC nodes (sorted, BCvelNodes of BCH:1549-1573) with their u,v,w
 * values (n*3), and the zero-pressure node (BCH:705-706; <= 0 disables the
 * pin, reference quirk BCH:4172). */
int fsx_upload_bc(fsx_ctx* ctx, int n, const int* nodes, const double* values, int zero_p_node);
/* NGP in {1, 8, 27}; density and viscosity as in BCH:559-560; dt BCH:551. */
int fsx_set_params(fsx_ctx* ctx, int NGP, double density, double viscosity, double dt);
/* pressure solver: mode (fsx_solver_mode), relative tolerance, iteration cap */
int fsx_set_solver(fsx_ctx* ctx, int mode, double rtol, int maxit);
/* Velocity mass of the step1 and step3 updates. FSX_MASS_LUMPED (default) is
 * the reference: UnpHalf = Un + dt R1 MdInv, Acc = R3 MdInv. With
 * FSX_MASS_CONSISTENT the updates solve M_bc a = R1 and M_bc Acc = R3
 * instead, M_bc being the assembled consistent mass with the rows and
 * columns of BC nodes replaced by the identity; the three velocity
 * components are solved together as one 3-right-hand-side Jacobi-PCG (one
 * application of M per iteration, per-component scalars and stop test
 * ||r_c|| <= rtol ||b_c||). Each component starts from the update's previous
 * solution (the last a, or Acc) when that leaves a smaller residual than 0,
 * else from 0 (0 after fsx_init_state). M is applied element by element,
 * y = sum_e M_e p_e from detJ w at the Gauss points, summed per node in a
 * fixed element order; FSX_MASS_CONSISTENT_CSR applies the assembled CSR
 * values instead (also what FSX_MASS_CONSISTENT does after M values are
 * uploaded, until the next fsx_assemble_static). The pressure operator Z
 * keeps the lumped MdOrigInv. No reference counterpart (the reference
 * assembles M, then lumps it, BCH:3195-3295). Needs the mesh uploaded. */
int fsx_set_mass(fsx_ctx* ctx, int mode, double rtol, int maxit);
/* Arithmetic of the consistent-mass solves (element form): 64 (default) or
 * 32, the mixed forms': FP32 Jacobi-PCG segments on the correction, each
 * starting from the FP64 residual b - M_bc x and ending at a 1e-4 reduction
 * of its FP32 residual, then x += d and the FP64 residual, whose
 * per-component test ||r_c|| <= rtol ||b_c|| decides convergence as in the
 * FP64 solve. The CSR form (FSX_MASS_CONSISTENT_CSR, or M values uploaded)
 * stays FP64. No reference counterpart (the reference lumps M). */
int fsx_set_mass_precision(fsx_ctx* ctx, int bits);
/* iterations (largest over the components is the count of the block
 * solve) and largest relative residual of the last mass solve, and the
 * iterations of all mass solves so far; any pointer may be NULL */
int fsx_mass_stats(fsx_ctx* ctx, int* last_iters, double* last_relres, long long* total_iters);
/* Storage precision of K, G and G^T as the step kernels read them (step1,
 * K*Acc, step2's right-hand side, step3): 64 (default, the reference) or 32,
 * the mixed-precision forms' FP32 assembly: FP32 copies rounded to nearest
 * from the FP64 assembly, products and sums still in FP64. Z, the pressure
 * solve's FP64 residual and every vector keep FP64. Costs 4 bytes per K
 * entry and 24 per G entry of device memory. No reference counterpart (the
 * reference is FP64 throughout; its `real` type is unused, BCH:130-134). */
int fsx_set_matrix_precision(fsx_ctx* ctx, int bits);
/* At matrix precision 32, whether step3 and step2's right-hand side read G
 * and G^T as row classes (1, the default) where the rows compress: rows of
 * G with the same column offsets from their first column and the same FP32
 * values (components below 2^-24 of the row's largest magnitude dropped)
 * share one class list, so each row costs its first column and a class
 * reference instead of 16 bytes an entry. G^T's rows share their values by
 * class too, and their column offsets where those repeat (the device
 * numbers the velocity nodes in grid order on structured meshes), else
 * they keep their columns (fsxh_g_classes builds the same classes on the
 * host). The form is taken where the class lists hold at most an eighth of
 * G's entries: on meshes of congruent elements (the uniform cavity: 64
 * classes of G and 27 of G^T, from 10^3 elements on); elsewhere (a
 * stretched or curved mesh) the kernels read the FP32 CSR copies. 0 keeps
 * the CSR copies. No reference counterpart (the reference stores G as three
 * CSR value arrays, BCH:3322-3513). */
int fsx_set_matrix_classes(fsx_ctx* ctx, int on);
/* Whether the learned solver's LDS-tiled Z products (there where Z's FP32
 * rows form classes on a structured grid) apply Z as Kronecker products of
 * 1D stencils (1, the default) where the pressure grid is a tensor-product
 * box grid and that form reproduces Z's rows to 1e-10 of each row's
 * largest magnitude (the pinned node's penalty added on its diagonal;
 * fsxh_z_tensor builds the same tables on the host): 45 multiply-adds a
 * row from three passes over the tile instead of the 81 entries of its
 * FP32 row class. 0 keeps the row classes. Rebuilds the solvers' forms of Z
 * when it is built. No reference counterpart (the reference applies the
 * assembled CSR Z, CU:418-451). */
int fsx_set_z_tensor(fsx_ctx* ctx, int on);
/* in_use: 1 when the tiles apply the Kronecker form; err: its largest
 * mismatch against Z's rows (-1 when the grid is not a box grid or Z is not
 * built); pointers may be NULL */
int fsx_z_tensor_info(fsx_ctx* ctx, int* in_use, double* err);
/* The class form in use: G's classes and their entries, G^T's classes and
 * theirs (all 0 when the kernels read the CSR copies); pointers may be NULL */
int fsx_g_class_info(fsx_ctx* ctx, int* classes, long long* entries, int* t_classes, long long* t_entries);
/* step1 with FP32 operand shadows on a structured mesh whose K patterns are
 * steps of at most 2 on the velocity grid can stage each block's column
 * window of shadow records in LDS (enable != 0) instead of gathering them
 * from memory (0, the default: the faster form at 64^3); the products,
 * their order and so the results are the same either way. No reference
 * counterpart (a kernel-form switch for measurement and checks). */
int fsx_set_step1_window(fsx_ctx* ctx, int enable);
/* step1 of the mixed forms element by element (enable != 0, the default):
 * when the convection and the matrices are FP32, the convection is in its
 * element form, the mass is lumped and K is the assembled one (not uploaded
 * through fsx_upload_vec), fsx_convection only records the convection and
 * the next fsx_step1 forms it together with K UnpHalf_prev and the pending
 * K Acc_prev from the element forms (FP32 matrix-core products of the shape
 * tables with the FP32 shadows of Un, UnpHalf_prev and Acc_prev, the node
 * sums FP64) instead of the CSR products of K; then the same update of R1,
 * UnpHalf and KAcc as the CSR form, equal to FP32 rounding. Any other entry
 * point that reads R1 or changes the state first forms a recorded
 * convection on its own. 0: the CSR products. No reference counterpart
 * (the reference applies the assembled K, BCH:3746-3783). */
int fsx_set_step1_elements(fsx_ctx* ctx, int enable);
/* Arithmetic of the matrix-free convection residual (fsx_convection): 64
 * (default, the reference's) or 32, the mixed forms' FP32 element assembly:
 * shape tables, node values, Jacobians and Gauss-point sums in FP32, the
 * element result subtracted from the FP64 R1. No reference counterpart. */
int fsx_set_convection_precision(fsx_ctx* ctx, int bits);
/* Form of the FP32 convection: elements != 0 (the default) one pass over all
 * elements, one thread per element and velocity component, the element rows
 * summed per node in a fixed order; 0 one pass per colour of the element
 * colouring with the rows added into R1 (the FP64 convection is always the
 * latter). Both compute each element's rows in the same FP32 arithmetic. No
 * reference counterpart (a kernel-form switch for measurement and checks). */
int fsx_set_convection_form(fsx_ctx* ctx, int elements);
/* Initial guess of the pressure solves (fsx_step2, fsx_solve_pressure).
 * FSX_X0_PREVIOUS starts from the Pdot already held (zero after
 * fsx_init_state, as the reference zeroes Pdot_d once): r0 = b - Z x0 in
 * FP64, the selected solver on Z e = r0 from zero with its relative
 * tolerance scaled by |b|/|r0|, x = x0 + e, so the stop test is the cold
 * solve's |b - Z x| <= rtol |b|; an x0 already within it takes no
 * iteration. FSX_X0_PROJECTION starts from the best approximation of the
 * solution, in the Z norm, in the span of the last 16 solutions (projection
 * for successive right-hand sides, after P. F. Fischer, CMAME 163 (1998)):
 * x0 = X a with (X^T Z X) a = X^T b, Z x_j kept from each solve's final
 * FP64 residual (taken as the b_j it was solved for in the FP64 solver,
 * which keeps no FP64 residual), vectors that add less than 1e-10 of their
 * own Z norm to the span dropped; then as FSX_X0_PREVIOUS from x0. The
 * window restarts when Z is built or the state initialised; a solve that
 * ends at maxit is not added. Same stop test; on the 64^3 cavity its x0
 * leaves a residual about 300 times smaller than the previous solution's
 * (32 times with 8 solutions). Not for the direct solver (it takes no
 * initial guess). */
int fsx_set_solver_x0(fsx_ctx* ctx, int policy);
/* Order and factor form of FSX_SOLVER_DIRECT (fsx_set_direct_order) */
enum fsx_direct_order {
  FSX_DIRECT_BAND = 0,  /* Z as a band in the narrower of the device's numbering and reverse
                           Cuthill-McKee, factored in place (no fill outside the band) */
  FSX_DIRECT_ND = 1     /* the default: a fill-reducing nested-dissection order (the role of the
                           prototype's symamd, ML:1880-1881) and a multifrontal factor, one dense
                           front per separator, the fronts of a tree level factored together */
};
/* Selects the direct solve's order and factor (FSX_DIRECT_ND by default);
 * a change takes effect at the next direct solve, which factors Z in the
 * new form. No counterpart in the reference's C++/CUDA code. */
int fsx_set_direct_order(fsx_ctx* ctx, int order);
/* FSX_SOLVER_DIRECT's factor: half-bandwidth of Z in the band order (-1
 * unless the band factor is the current one), wall time of the current
 * factorization (ms) and the bytes of its factor (the band, or L's columns
 * of every front). The factor is formed at the first direct solve after Z
 * changes (ML:1876-1885 factors once before the time loop); the band must
 * fit in 64 GiB and the fronts in the free device memory (FSX_ENOMEM
 * otherwise). No counterpart in the reference's C++/CUDA code. */
int fsx_direct_stats(fsx_ctx* ctx, int* half_bandwidth, double* factor_ms, double* band_bytes);
/* The nested-dissection factor's shape (zeros and depth -1 before one is
 * formed): bytes of L's front columns, factor flops (sum of squared column
 * counts, as fsxh_nd_order reports them), number of fronts and depth of the
 * separator tree. No reference counterpart. */
int fsx_direct_info(fsx_ctx* ctx, double* factor_bytes, double* factor_flops, int* fronts, int* depth);
/* enqueue solver iterations as replayed HIP graphs (1) or as individual
 * launches (0, the default: on the 64^3 cavity both run within 5% of each
 * other); results are identical. No reference counterpart (the CUSP and MKL
 * solvers are opaque library calls, CU:418-451). */
int fsx_set_graphs(fsx_ctx* ctx, int enable);

/* ---------------------------------------------------------------- assembly */
/* step0 values (BCH:3127-3315): K = nu*sum grad.grad, M, G_d = -(1/rho) sum
 * Sp d_d Sv, lumped Md, MdOrigInv, MdInv with BC rows 1 (BCH:4121-4135). */
int fsx_assemble_static(fsx_ctx* ctx);
/* Z = sum_d G_d^T diag(MdOrigInv) G_d on the uploaded Z pattern, penalty
 * x1000 at the zero-pressure node; replaces calculateZ_CUSP (CU:281-411). */
int fsx_build_Z(fsx_ctx* ctx);
/* zero state + initial velocity BCs: initializeAndAllocate (BCH:2718-2808)
 * and applyBC_initial (BCH:4062-4107) */
int fsx_init_state(fsx_ctx* ctx);

/* ---------------------------------------------------------------- time step
 * The driver calls these in the order of timeLoop (BCH:2859-3118). */
/* UnpHalf_prev <- Un, Pnp1_prev <- Pn, Acc_prev <- 0   (BCH:2865-2881) */
int fsx_begin_step(fsx_ctx* ctx);
/* R1 <- -A(Un) * UnpHalf_prev, matrix free; replaces calculateMatrixAGPU
 * (CU:906-939) / calculateMatrixA (BCH:3520-3685) */
int fsx_convection(fsx_ctx* ctx);
/* R1 -= K UnpHalf_prev + G Pn; BC rows 0; UnpHalf = Un + dt R1 MdInv.
 * replaces step1GPUpart (CU:946-973) */
int fsx_step1(fsx_ctx* ctx, int iter);
/* R2 = G^T (UnpHalf/dt^2 - [iter>1] MdOrigInv KAcc); R2[p0] = 0; solve
 * Z Pdot = R2; Pnp1 = Pn + dt Pdot. replaces step2GPU (CU:980-1022) and
 * CUSP_CG_solver (CU:418-451). pcg_iters/relres may be NULL. Returns
 * FSX_ENOCONV if the solver stopped at maxit (the solution is still used,
 * as the reference does). */
int fsx_step2(fsx_ctx* ctx, int iter, int* pcg_iters, double* relres);
/* R3 = -dt G Pdot - [iter>1] dt KAcc; BC rows 0; Acc = R3 MdInv;
 * Unp1 = UnpHalf + dt Acc. replaces step3GPU (CU:1029-1060) */
int fsx_step3(fsx_ctx* ctx, int iter);
/* ||Unp1-Unp1_prev||/||Unp1|| and ||Pnp1-Pnp1_prev||/||Pnp1||, converged =
 * both < tol (NaN counts as not converged). replaces checkConvergenceGPU
 * (CU:1084-1121) */
int fsx_check_convergence(fsx_ctx* ctx, double tol, int* converged, double* rel_u, double* rel_p);
/* UnpHalf_prev <- UnpHalf, Unp1_prev <- Unp1, Acc_prev <- Acc,
 * Pnp1_prev <- Pnp1 (BCH:2964-2972, CU via BCH:2925-2932); Unp1 and Pnp1
 * keep their values (they become Un, Pn after the last iteration). */
int fsx_next_iteration(fsx_ctx* ctx);
/* KAcc = K Acc_prev; replaces calculate_KtimesAcc_prevGPU (CU:1067-1077).
 * The product is formed by the next fsx_step1 in the same pass over K as
 * K UnpHalf_prev; any other call that reads KAcc or changes Acc_prev first
 * forms it on its own, so the result is the same at every observable point. */
int fsx_kacc(fsx_ctx* ctx);
/* maxAcc = max |Unp1 - Un| / dt (BCH:3048-3062; the reference GPU variant
 * omits the /dt, CU:99-118), steady = maxAcc <= converge; then Un <- Unp1,
 * Pn <- Pnp1 (BCH:3070-3077). replaces checkConvergenceInTimeGPU
 * (CU:1151-1178) */
int fsx_end_step(fsx_ctx* ctx, double converge, int* steady, double* max_acc);
/* the four monitor values u,v,w,p of Un/Pn at one node, without copying the
 * whole field (replaces the full D2H of printMonitorDataGPU, CU:1128-1144) */
int fsx_monitor(fsx_ctx* ctx, int node, double* out4);

/* ---------------------------------------------------------------- data */
int fsx_download(fsx_ctx* ctx, int id, double* host, size_t count);
/* State vectors and K, M, Z values (not G or the lumped masses). A K upload
 * refreshes the FP32 copy the step kernels read; an M upload lumps it again
 * (Md, MdOrigInv, MdInv: BCH:3262-3295; Z keeps the lumped mass it was built
 * from until fsx_build_Z runs again); a Z upload (values with the penalty
 * pin, as fsx_download returns them) rebuilds every form the pressure
 * solvers read and makes Z current without fsx_build_Z. */
int fsx_upload_vec(fsx_ctx* ctx, int id, const double* host, size_t count);
int fsx_sync(fsx_ctx* ctx);
/* solve Z x = b on the context's Z with the current solver settings; b and
 * x are host arrays of NNp doubles (a test and tooling entry point). */
int fsx_solve_pressure(fsx_ctx* ctx, const double* b, double* x, int* iters, double* relres);
/* The FP64 Jacobi-PCG of fsx_solve_pressure (the plain solver, x0 = 0, the
 * context's rtol and maxit, whatever solver mode and x0 policy are set) on
 * DEVICE arrays b and x of NNp doubles, allocating nothing: its solve
 * scalars, reduction partials and Krylov vectors live in a file-scope device
 * array of the library (one per GPU, held until the process ends), so the
 * call makes no hipMalloc/hipFree and no allocating library call. x equals
 * what fsx_solve_pressure returns in that configuration. Limit: NNp <= 2^20
 * pressure nodes (about 32 MiB of scratch); larger Z is refused with
 * FSX_EINVAL. Every call on one GPU shares that array: calls that overlap on
 * one GPU (from two contexts or threads) race; nothing of a previous call is
 * read. b and x must be ready for, and are read back after, the context's
 * stream (the call returns once the solve has finished). No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under stream capture, tight serving loops, harnesses that forbid it). */
int fsx_solve_pressure_noalloc(fsx_ctx* ctx, const double* b, double* x, int* iters, double* relres);
/* y = alpha * op(v) + beta * y on the context's matrices for host vectors:
 * which = 0 K (NN), 1..3 G_d (NN <- NNp), 4 Z (NNp), 5..7 G_d^T (NNp <- NN).
 * Mirrors the reference's mkl_dcsrmv / cusparseDcsrmv calls. which = 8 and
 * 9 apply Z as the FP64 and the FP32 solvers read it (structural zeros
 * dropped; FP32 rounds v, Z and y) with alpha = 1, beta = 0 only. */
int fsx_spmv(fsx_ctx* ctx, int which, const double* x, double* y, double alpha, double beta);

/* Times `reps` back-to-back launches of the pressure solver's SpMV kernel
 * (q = Z p) on the context's stream with HIP events. precision 0 = FP64 Z,
 * 1 = FP32 Z (the mixed-precision and learned inner solves). Returns the
 * mean device time per launch in ms and the algorithmic HBM bytes of one
 * launch in the form the solvers read: FP64 values + column indices (or the
 * row-pattern ids and dictionary) + row pointers; or, for FP32 row classes,
 * the launch plan (a 16-byte record per chunk of rows of one class and a
 * 4-byte row index per row) + the class table; plus one read of p and one
 * write of q. No reference counterpart (the reference prints wall times only,
 * BCH:4489-4518). */
int fsx_bench_spmv(fsx_ctx* ctx, int precision, int reps, double* ms_per_launch, double* bytes_per_launch);
/* Size of the dictionary of row patterns (distinct column-offset lists) of
 * the solvers' Z (structural zeros dropped, fsx_z_forms) that the Z SpMVs
 * read instead of a column index per entry; 0 patterns when the rows did not
 * compress and the SpMVs read CSR columns. No reference
 * counterpart (the reference's Z is plain CSR, CU:281-411). */
int fsx_z_patterns(fsx_ctx* ctx, int* npatterns, long long* offsets);
/* The solvers' forms of the built Z. Entries whose magnitude is below 2^-46
 * of their row's largest are structural zeros (couplings that vanish
 * analytically, left by G^T M^-1 G as cancellation residue near 1e-37) and
 * are dropped; nnz_solver returns the entries kept. The FP32 copy is stored
 * as row classes when its rows take at most 65535 distinct lists of (column
 * offset, FP32 value) with at most 2^20 entries in total: nclasses and
 * class_entries return their count and total length (0 when the FP32 solves
 * read the FP32 rounding of the kept entries on their CSR rows). No reference counterpart (the reference solves
 * with the assembled CSR Z, CU:418-451, BCH:3981-4055). */
int fsx_z_forms(fsx_ctx* ctx, long long* nnz_solver, int* nclasses, long long* class_entries);
/* Times `reps` launches of a plain device copy of `bytes` (read once, written
 * once) between two scratch buffers: the HBM bandwidth reference and the
 * known-traffic kernel that calibrates PMC byte counters. No reference
 * counterpart. */
int fsx_bench_copy(fsx_ctx* ctx, size_t bytes, int reps, double* ms_per_launch);

/* Row kernels of the time step for fsx_bench_kernel */
enum fsx_kernel_id {
  FSX_KERNEL_STEP1 = 0,      /* K 3-RHS + G 3-block SpMV with the step1 update (step1GPUpart, CU:946-973) */
  FSX_KERNEL_KACC = 1,       /* K 3-RHS SpMV (calculate_KtimesAcc_prevGPU, CU:1067-1077) */
  FSX_KERNEL_STEP2_RHS = 2,  /* G^T 3-block SpMV forming R2 (step2GPU, CU:980-1011) */
  FSX_KERNEL_STEP3 = 3,      /* G 3-block SpMV with the step3 update (step3GPU, CU:1029-1060) */
  /* step1 of the mixed forms element by element (fsx_set_step1_elements):
   * its element pass (convection, K UnpHalf_prev and K Acc_prev of every
   * element on the matrix cores) and its node pass (the element rows summed
   * per node, step1's update); replace step1GPUpart's K and G products,
   * CU:946-973 */
  FSX_KERNEL_STEP1_ELEMENTS = 4,
  FSX_KERNEL_STEP1_NODES = 5
};
/* Times `reps` back-to-back launches of one of those kernels (inner
 * iteration >= 2 form, with KAcc) on the context's stream with HIP events:
 * mean device time per launch in ms, and its algorithmic HBM bytes (every
 * matrix entry, row pointer and vector element it touches counted once;
 * FP64 values, int32 indices, K's columns as the kernels read them: one
 * pattern start per row plus the pattern dictionary when K's rows compress,
 * else a column per entry). Clobbers the step vectors: call
 * fsx_init_state before stepping again. No reference counterpart (the
 * reference prints wall times only, BCH:4489-4518). */
int fsx_bench_kernel(fsx_ctx* ctx, int kernel, int reps, double* ms_per_launch, double* bytes_per_launch);

/* kernel time accounting (device time of the last timed region, ms) */
typedef struct fsx_timing {
  double pcg_ms;         /* pressure solve kernels */
  double pcg_spmv_ms;    /* of which the Z SpMV kernel */
  long long pcg_spmv_launches;
  double other_ms;       /* everything else of a step */
  long long pcg_iters;
} fsx_timing;
int fsx_timing_reset(fsx_ctx* ctx, int enable);
int fsx_timing_get(fsx_ctx* ctx, fsx_timing* out);

/* ---------------------------------------------------------------- learned preconditioner
 * No counterpart in the reference (its only preconditioner is Jacobi,
 * CU:418-451, BCH:3981-4055). A 3-level linear 3-D conv U-net (channels
 * 16/32/64, 3x3x3 kernels, BF16 weights, FP32 accumulation) acting on the
 * pressure grid; FSX_SOLVER_LEARNED runs flexible PCG with
 * z = dinv r + U(r), falling back to Jacobi-PCG when the flexible recurrence
 * stalls. The pressure nodes must be numbered as a structured nx*ny*nz grid,
 * x fastest (the corner numbering of the structured generators), with
 * nx, ny, nz = 8m+1 (m >= 1: three node-aligned coarsenings, 65 -> 33 -> 17
 * -> 9 nodes per axis on the 64^3 cavity). The network is
 * J^T J with J = three stride-2 3x3x3 convolutions C0 (1 -> 16 channels),
 * C1 (16 -> 32), C2 (32 -> 64), each followed by zeroing the boundary nodes
 * of its output grid; the transpose pass reuses the same weights. */
int fsx_unet_setup(fsx_ctx* ctx, int nx, int ny, int nz, unsigned long long seed);
/* number of FP32 parameters (70144 = 16*32 + 32*448 + 64*864): the three
 * layers in order C0, C1, C2, each stored as weights [cout][kp] with
 * kp = 27*cin rounded up to 32 (C0 32, C1 448, C2 864), entry k = tap*cin + ci
 * of row co, tap = (dx+1) + 3(dy+1) + 9(dz+1) for the input node
 * 2o + (dx, dy, dz) of output node o; C0 starts at offset 0, C1 at 512, C2 at
 * 14848; padding entries (k >= 27*cin) are 0. The weights are grid-size
 * independent: a set trained on one 8m+1 grid applies to any other. */
int fsx_unet_params(fsx_ctx* ctx, long long* n);
int fsx_unet_get_weights(fsx_ctx* ctx, float* host, size_t n);
int fsx_unet_set_weights(fsx_ctx* ctx, const float* host, size_t n);
/* output gain g of y = g * s * U(x / s): set if set > 0, returned in *get */
int fsx_unet_gain(fsx_ctx* ctx, float set, float* get);
/* y = U(x) (s = 1) for host arrays of NNp floats */
int fsx_unet_apply(fsx_ctx* ctx, const float* x, float* y);
/* loss of y = gain * U(x) against t and its parameter gradient (nparam
 * floats), all on the FP32 Z: kind 0 relative L2 |y - t|^2 / |t|^2, kind 1
 * relative energy norm (y - t)' Z (y - t) / t' Z t (the additive form's
 * training loss), kind 2 the energy norm after post-smoothing,
 * f' Z f / t' Z t with f = (I - S Z)(y - t) and S = omega D^-1 of the
 * two-level form (its training loss) */
int fsx_unet_loss_grad(fsx_ctx* ctx, int kind, const float* x, const float* t, double* loss, float* grad);
/* keep the right-hand sides and solutions of the next `capacity` converged
 * pressure solves (0 stops capturing); a call with capacity > 0 drops what an
 * earlier capture kept */
int fsx_unet_capture(fsx_ctx* ctx, int capacity);
int fsx_unet_captured(fsx_ctx* ctx, int* n);
/* training pairs from the captured solves, one per truncation depth ks[q]:
 * the residual r of the ks[q]-th iterate from x = 0 of FP64 Jacobi-PCG (flags
 * bit 0 clear) or of the learned flexible PCG (bit 0 set) and the error e
 * that remains, both scaled by 1 / rms(r). In the additive form the input is
 * r and the target e - dinv r; in the two-level form the input is
 * r1 = r - Z S r and the target e - S r (the error after pre-smoothing).
 * Bit 1 appends to the pairs already held (of the same form), otherwise they
 * are replaced. */
int fsx_unet_prepare(fsx_ctx* ctx, const int* ks, int nk, int flags, int* nsamples);
/* `count` more training pairs, of the form in use, without solves: seeded
 * errors e = (I - S Z)^sweeps e0 with e0 i.i.d. normal per node (0 at the
 * pinned node), S = omega D^-1, and r = Z e; appended to the pairs already
 * held. Drawn errors cover every mode of Z, so that training does not trade
 * modes absent from the captured solves for the ones present. */
int fsx_unet_prepare_random(fsx_ctx* ctx, int count, int sweeps, unsigned long long seed, int* nsamples);
/* Adam on seeded mini-batches of the prepared pairs, normalised per pair by
 * its error energy e' Z e: in the additive form the energy norm
 * (y - t)' Z (y - t), in the two-level form the energy f' Z f of the error
 * f = (I - S Z)(t - y) that the whole preconditioner leaves (kinds 1 and 2
 * of fsx_unet_loss_grad). Returns the mean loss of the last 50 steps (the
 * fraction of error energy that one application leaves). */
int fsx_unet_train(fsx_ctx* ctx, int steps, int batch, float lr, unsigned long long seed, double* loss);
/* Form of the preconditioner M that the flexible PCG applies:
 * form 0, additive: z = dinv r + gain J^T J r;
 * form 1, two-level (the default): z = S r + C r1 + S (r - Z (S r + C r1)),
 * r1 = r - Z S r, with S = omega D^-1 (one Jacobi sweep) and C = gain J^T J
 * the network as the coarse correction, so I - M Z = (I - S Z)(I - C Z)(I - S Z)
 * and M is SPD for any weights while omega < 2 / lambda_max(D^-1 Z) (about 2
 * on the cavity meshes: omega = 0.8 by default). Form 1 needs a structured
 * pressure grid whose Z rows compress into classes (its three Z products read
 * LDS-staged tiles of the grid); elsewhere the solver uses form 0. Training
 * pairs (fsx_unet_prepare) follow the form set when they are made. omega is
 * taken for form 1 only and must lie in (0, 1). No reference counterpart. */
int fsx_unet_form(fsx_ctx* ctx, int form, float omega);
/* the form the learned solver runs (1 only where the tiles exist) and omega */
int fsx_unet_form_get(fsx_ctx* ctx, int* form_in_use, float* omega);
/* z = M r for host arrays of NNp floats, with the form in use (a test and
 * tooling entry point; overwrites the solver's work vectors) */
int fsx_unet_precondition(fsx_ctx* ctx, const float* r, float* z);
/* flexible-PCG iterations allowed before the Jacobi-PCG fallback (default 400) */
int fsx_unet_limit(fsx_ctx* ctx, int learned_maxit);
/* flexible-PCG iterations of the last learned solve and why it fell back to
 * Jacobi-PCG: 0 it did not, 2 iteration cap, 6 loss of descent */
int fsx_unet_stats(fsx_ctx* ctx, int* last_learned_iters, int* last_fallback);

#ifdef __cplusplus
}
#endif
#endif /* FSX_H */
