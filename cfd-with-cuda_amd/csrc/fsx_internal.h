// Internal definitions shared by the HIP translation units of libfsx.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/fsx.h"

namespace fsx {
struct MfState;   // fsx_sparse.hip

constexpr int kWave = 64;         // CDNA wavefront width
constexpr int kBlock = 256;       // default block size for row kernels
constexpr int kMaxGrid = 2048;    // row kernels: 256 CUs x 8 blocks, a multiple of 8 XCDs
constexpr int kZRowMax = 1024;    // longest Z row handled by the Z-build kernel
// Entries of Z whose magnitude is below this fraction (2^-46) of their row's
// largest are structural zeros: couplings that vanish analytically and that
// the triple product G^T M^-1 G leaves as cancellation residue (about 1e-37
// on the cavity meshes, against 1.6e-6 for the smallest coupling that does
// not vanish). The solvers' copies of Z drop them; the assembled Z keeps them.
constexpr double kPruneRel = 1.4210854715202004e-14;

// Device-resident scalars of one Krylov solve, written by block 0 of a kernel
// and read by later launches. Values that a kernel both reads (all blocks)
// and updates (block 0) are versioned by the parity of the iteration index j
// that the host passes in, so no block can see a value of the same launch.
struct SolveScalars {
  double rz_v[2];     // r.z entering iteration j sits in rz_v[j & 1]
  double rr, bb, rtol2bb;
  double rr_anchor;   // mixed solver: FP64 residual^2 at the last reliable update
  double rr_hi;       // mixed solver: latest FP64 residual^2
  int iters;          // completed iterations
  int stop;           // 0 running, 1 converged, 2 maxit, 3 NaN/Inf, 5 paused for a reliable update
  int maxit, pad;
  int active_v[2];    // iteration j ran its update kernel (stop was 0 then)
  double rp_v[2];     // learned solver: r.p entering iteration j sits in rp_v[j & 1]
  double pq_v[2];     // learned solver: p.Zp of iteration j
  double rp;          // tiled learned solver: r.p of the current iteration (written by its direction kernel)
  double pq_prev;     // tiled learned solver: p.Zp of the last update (0 before the first)
};

// SolveScalars as published to host-coherent pinned memory by k_publish:
// the host spins on seq instead of a copy plus a stream synchronisation.
// FSX_X0_PROJECTION: solutions kept, and the dot products one guess forms
// (X^T b and one kept slot's row and column of the Gram matrix)
constexpr int kProjMax = 16;
constexpr int kProjDots = 3 * kProjMax;

struct Published {
  SolveScalars s;
  unsigned seq;
  unsigned pad[3];
  double proj[kProjDots];   // the projection start's dot products (k_proj_publish)
};

// The pressure operator Z staged through LDS on a structured nx*ny*nz grid
// (x fastest): a tile is every x of ty rows in y and tz in z (at most 1024
// rows, one a thread), held in LDS with a halo of 2 nodes ((nx + 4) x (ty +
// 4) x (tz + 4) values, zero outside the grid), since Z couples nodes up to 2
// apart per axis. A tile's rows are sorted by FP32 row class; row t of tile b
// is rows[b * rstride + t] = (grid node, LDS index | own index << 16, global
// class entry start, entry count | tile entry start << 8), its entries
// (LDS offset, FP32 value bits) sit at ent + global start and, for the
// wavefronts whose rows mix classes, in the tile's entry block
// tent[b * estride ...] that the kernel stages in LDS.
//
// On a tensor-product box grid the tiles apply Z as Kronecker products of 1D
// stencils instead (fsx_ztensor.h): `coef` holds per axis and grid point
// K's five coefficients (offsets -2..2) then M's, FP32, the x axis first,
// and the tile's rows are in grid order; `zp` (-1: none) is the pinned
// node, whose diagonal takes `pen` beyond the tensor form.
struct TilePlan {
  int nx, ny, nz;
  int ty, tz, lx, ly, lz;
  int tiles_y, ntiles, npart;
  int rstride, estride;    // rows and entries reserved per tile
  const int4* rows;
  const int2* ent;
  const int2* tent;
  int tensor;              // 1: the Kronecker form (coef, zp, pen), 0: the row classes
  const float* coef;
  int zp;
  float pen;
};

// Vectors of one tiled kernel of the learned solver's iteration (the mode
// decides which are read and written; see k_tile in fsx_stencil.hip).
struct TileVecs {
  const float* z;     // kTileDir: the preconditioned residual
  const float* pin;   // kTileDir: the last direction; kTileSmooth, kTilePre: the direction for the r.p partial
  float* pout;        // kTileDir: the new direction
  float* q;           // kTileDir writes Z p; kTileSmooth and kTilePost read it
  const float* rin;   // kTileSmooth: the residual before the update; kTilePre, kTilePost: the residual
  float* rout;        // kTileSmooth: the updated residual
  float* x;           // kTileSmooth: the FP32 iterate
  float* r1;          // kTileSmooth, kTilePre: r - Z (omega dinv r), the network's input
  const float* z2;    // kTilePost: omega dinv r + the network's correction
  float* zout;        // kTilePost: z
  const float* dinv;  // FP32 Jacobi inverse
  float omega;        // Jacobi smoothing weight
  double* part;       // kTileSlots slots of TilePlan::npart partial sums
  SolveScalars* sc;
  int j, maxit;       // iteration index (kTileDir) and flexible-iteration cap
  double delta2;      // reliable-update threshold on |r|^2 / |r_anchor|^2
};
enum { kTileDir = 0, kTileSmooth = 1, kTilePre = 2, kTilePost = 3 };
constexpr int kTileSlots = 5;   // p.q, r.r, r.p, r.z, z.q

// Scalars of the time step (norms for the convergence checks, max |dU|/dt).
struct StepScalars {
  double uu, ud, pp, pd, maxacc;
};

// Shape-function tables on the device, GP-major (k = Gauss point):
//   Sv[k*27 + i], dSv[(k*27 + i)*3 + m], Sp[k*8 + j], dSp[(k*8 + j)*3 + m], w[k]
struct ShapeTables {
  double* Sv;
  double* dSv;
  double* Sp;
  double* dSp;
  double* w;
};
// the same tables rounded to FP32 (the FP32 element kernels read them as
// scalar loads: every thread of a wavefront walks the same Gauss point)
struct ShapeTablesF {
  const float* Sv;
  const float* dSv;
  const float* Sp;
  const float* dSp;
  const float* w;
};

// A block of solver iterations captured once as a HIP graph (one per
// starting parity; kernels see the iteration index only through j & 1) and
// replayed, since host launch cost otherwise dominates the short vector
// kernels of an iteration. `key` names what the captured arguments depend
// on beyond the context's own buffers; a new key recaptures.
struct IterGraph {
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  unsigned long long key = 0;
};
constexpr int kGraphBlock = 2;   // iterations per captured graph (even)

// Iterations expected before the next host decision (a reliable update every
// delta of residual reduction, or convergence at rtol), from the previous
// solve's count: chunks sized so, then topped up in small steps, keep the
// no-op launches after a stop few.
inline int segment_guess(int last_iters, double delta, double rtol, int floor_) {
  if (last_iters <= 0) return floor_;
  double f = 1.0;
  if (delta > 0.0 && delta < 1.0 && rtol > 0.0 && rtol < delta) f = std::log(delta) / std::log(rtol);
  const int g = (int)(0.9 * last_iters * f);
  return g > floor_ ? g : floor_;
}

}  // namespace fsx

struct fsx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // sizes and parameters
  int NE = 0, NN = 0, NNp = 0, NGP = 27;
  double density = 1.0, viscosity = 0.01, dt = 0.01;
  int solver_mode = FSX_SOLVER_FP64_JACOBI;
  double rtol = 1e-8;
  int maxit = 5000;
  bool have_mesh = false, have_M = false, have_G = false, have_Z = false, have_bc = false, have_params = false;
  bool assembled = false, z_built = false, state_ready = false;
  bool kacc_pending = false;     // fsx_kacc deferred to the next step1 (settle_deferred)
  bool gpn_valid = false;        // GPn holds G Pn for the current Pn
  double* GPn = nullptr;         // 3NN: G Pn, node-major, kept over the inner iterations of a step
  double* GPn_next = nullptr;    // 3NN: G Pnp1 from step3 (G Pn + dt G Pdot), the next step's G Pn
  bool gpn_next_valid = false;   // GPn_next holds G Pnp1 for the current Pnp1
  // velocity mass (fsx_set_mass): lumped (the reference) or consistent, M_bc
  // solved by the 3-right-hand-side Jacobi-PCG of fsx_mass.hip
  int mass_mode = FSX_MASS_LUMPED;
  double mass_rtol = 1e-10;
  int mass_maxit = 1000;
  double* mass_w = nullptr;      // r, z, p, q (4 x 3NN)
  double* mass_x = nullptr;      // 3NN, step1's solution
  double* mass_b = nullptr;      // 3NN, step3's right-hand side R3
  double* mass_dinv = nullptr;   // NN, 1 / M_ii, 0 at BC nodes
  double* mass_sc = nullptr;     // device MassScalars
  bool mass_dinv_ready = false;
  int mass_bits = 64;            // 32: FP32 segments with FP64 residuals (element form)
  float* mass_w32 = nullptr;     // r, z, p, q, d (5 x 3NN) of the FP32 segments
  float* mass_fac32 = nullptr;   // FP32 detJ w per element and Gauss point
  float* mass_ye32 = nullptr;    // FP32 element products
  float* mass_dinv32 = nullptr;  // FP32 Jacobi inverse
  bool mass32_ready = false;     // the FP32 copies match mass_fac and mass_dinv
  int mass_seg_iters[4] = {0, 0, 0, 0};   // iterations of the last mixed solve's segments
  // The element-by-element products (the consistent mass, the FP32
  // convection) visit the elements sorted by their centre node's internal id
  // (the nodes' coordinate order), so that a block's elements and a node's
  // elements sit close together in memory: el_L27 holds their velocity ids
  // transposed ([i][e]), el_P8 their corners' pressure ids ([e][m]), and the
  // node -> (element, local node) incidence lists the entries i * NE + e of
  // each node in ascending e. Built at the first element-form product.
  int* el_L27 = nullptr;
  int* el_P8 = nullptr;
  int* el_inc_ptr = nullptr;
  int* el_inc = nullptr;
  bool el_ready = false;
  // The consistent mass's detJ w per element and Gauss point ([k][e]) and its
  // per-element products ([i][e][3]). M values uploaded by fsx_upload_vec no
  // longer match the elements: until the next fsx_assemble_static the solves
  // then apply the CSR values.
  double* mass_fac = nullptr;
  double* mass_ye = nullptr;
  bool mass_fac_ready = false;
  bool mass_m_uploaded = false;
  // the FP32 convection's per-element products ([i][e][3]), and whether the
  // FP32 convection runs element by element (else per colour)
  float* conv_ye = nullptr;
  bool conv_elements = true;
  int mass_last_iters = 0;
  double mass_last_relres = 0.0;
  long long mass_total_iters = 0;
  // Internal velocity numbering. The device numbers the velocity nodes in
  // the order of their coordinates (z, then y, then x), so that the nodes a
  // row of K or G couples to sit close together in memory; the reference
  // numbers corners, mid-edge, mid-face and centre nodes in four separate
  // ranges (BCH:954-1320), which spreads every row's gathers over four
  // distant regions. vperm maps a reference id to the internal one, viperm
  // back. Everything at the boundary (patterns, vectors, values, BC nodes)
  // stays in the reference numbering and is converted on upload/download.
  std::vector<int> vperm, viperm;
  // the uploaded (reference-numbered) M and G patterns, for value downloads
  std::vector<int> refMrow, refMcol, refGrow, refGcol;
  std::vector<int> posM, posG;   // reference CSR position -> internal one (built on first use)
  // mesh (device)
  int* L27 = nullptr;            // NE*27, internal velocity ids
  int* P8 = nullptr;             // NE*8, pressure ids of the corner nodes (the reference's LtoGnode[e][0..7])
  double* xc = nullptr;          // NNp*3 corner coordinates
  int* color_order = nullptr;    // NE
  std::vector<int> color_ptr;    // host copy, ncolors+1
  // CSR patterns and values (device)
  long long nnzM = 0, nnzG = 0, nnzZ = 0;
  int* Mrow = nullptr; int* Mcol = nullptr;
  double* K = nullptr; double* M = nullptr;
  // FP32 copies of K, G, G^T read by the step kernels when the matrix
  // precision is 32 (fsx_set_matrix_precision); null at precision 64
  float* Kf = nullptr; float* Gf = nullptr; float* GTf = nullptr;
  float *Kf_mem = nullptr, *Gf_mem = nullptr, *GTf_mem = nullptr;   // their allocations
  // G and G^T as row classes (fsx_gclass.h) beside the FP32 copies, where the
  // rows compress and fsx_set_matrix_classes allows them; null otherwise.
  // step3 and step2's right-hand side read these instead of Gf / GTf.
  int2* Gc_row = nullptr;      // per row of G: (first column, class entry start << 8 | length)
  float4* Gc_ent = nullptr;    // per class entry: (column offset bits, G1, G2, G3)
  int* GTc_info = nullptr;     // per row of G^T: class entry start << 8 | length
  int* GTc_col = nullptr;      // G^T's columns in class order, row p from GTrow[p] (null with GTc_base)
  int* GTc_base = nullptr;     // per row of G^T: its first column, where the classes carry the offsets
  float4* GTc_val = nullptr;   // per class entry: (G1, G2, G3, column offset bits or 0)
  int gc_ncls = 0, gtc_ncls = 0;
  long long gc_nent = 0, gtc_nent = 0;
  bool use_g_classes = true;
  bool use_z_tensor = true;    // the tiles' Kronecker form of Z where it holds (fsx_set_z_tensor)
  // Z's Kronecker form in FP64 for the mixed and learned solvers' FP64
  // residuals (per axis and grid point K's then M's five coefficients, the
  // x axis first), null where the form does not hold
  double* zt64 = nullptr;
  int zt_n[3] = {0, 0, 0};
  int zt_zp = -1;
  double zt_pen = 0.0;
  double z_tensor_err = -1.0;  // its largest mismatch against Z's rows (-1: the grid is not a box grid)
  bool gc_built = false;       // the classes (or their absence) are of the current assembly of G
  int mat_bits = 64;
  int conv_bits = 64;            // arithmetic of the convection kernel (fsx_set_convection_precision)
  // FP32 shadows of UnpHalf and Acc (and of their "prev" buffers), interleaved
  // 8 floats a node (UnpHalf's 3 at Urec + 8i, Acc's at Arec = Urec + 4; the
  // 4th of each unused), written beside the FP64 vectors by step1 and
  // step3 at matrix precision 32 with lumped mass. step1 gathers its K
  // operands from the prev shadows while their flags say they hold the FP64
  // prev vectors rounded; any other writer of those vectors clears the flags.
  float *Urec = nullptr, *Urec_prev = nullptr, *Arec = nullptr, *Arec_prev = nullptr;
  float* rec_mem = nullptr;
  // direct pressure solve (FSX_SOLVER_DIRECT): Cholesky factor of Z as a
  // lower band in band order, entry (i, j) at i*band_W + (j - i + band_w)
  double* band = nullptr;
  double* band_up = nullptr;     // its transpose as an upper band, entry (j, i) at j*band_W + (i - j)
  double* band_inv = nullptr;    // inverses of the 1024x1024 diagonal blocks, then their transposes
  double* band_vec = nullptr;    // 3 band_n: the right-hand side, y and x, in band order
  double* band_part = nullptr;   // partial sums of the triangular-solve steps
  int* band_perm = nullptr;      // band order -> pressure id
  int* band_iperm = nullptr;     // pressure id -> band order
  int* band_flag = nullptr;      // set at a non-positive pivot
  long long band_W = 0;          // row stride
  int band_w = -1, band_n = 0;   // half-bandwidth, order padded to a multiple of 1024
  int band_nn = 0;               // NNp the permutations are sized for
  bool band_ready = false;       // the factor is of the current Z
  double band_factor_ms = 0.0;
  bool urec_ok = false, urec_prev_ok = false, arec_ok = false, arec_prev_ok = false;
  // step2's operand d = UnpHalf / dt^2 - [KAcc formed] MdOrigInv KAcc as an
  // FP32 shadow (same layout), written by step1 next to UnpHalf; valid while
  // drec_ok, with the KAcc term iff drec_kacc
  float* Drec = nullptr;
  bool drec_ok = false, drec_kacc = false;
  // compact FP32 shadow of Un (4 floats a node), the first inner iteration's
  // K operand (UnpHalf_prev = Un then); valid until the next inner iteration
  float* Ucmp = nullptr;
  bool ucmp_ok = false;
  bool ucmp_un = false;          // Ucmp holds Un rounded (set with the step's shadows, cleared when Un changes)
  // step1 element by element on the matrix cores (fsx_elem.hip): the
  // convection, K UnpHalf_prev and K Acc_prev from the element forms, then a
  // node pass. fsx_convection only records the convection (conv_pending)
  // when the next step1 can take this form; anything else that reads R1 or
  // changes the state first forms it on its own (settle_deferred).
  bool use_step1_elements = true;  // fsx_set_step1_elements
  bool conv_pending = false;
  float* ek_tab = nullptr;       // packed MFMA A operands per Gauss-point tile, then dSp and w per point
  int ek_ntiles = 0;
  int ek_grid = 0;               // blocks of the element kernel (0: 8 per CU)
  float* ek_ye = nullptr;        // per-element rows [i][e][6]: conv + K UnpHalf_prev, K Acc_prev
  float* ek_xl = nullptr;        // corner coordinates relative to corner 0, [e][8][3] (sorted elements)
  bool ek_xl_ready = false;
  double k_viscosity = 0.0;      // the viscosity K was assembled with
  bool k_uploaded = false;       // K values from fsx_upload_vec (no element form)
  // M/K columns as a dictionary of row patterns (as for Z below); null when
  // the rows do not compress
  int* Mrowpat = nullptr; int* Mpatoff = nullptr;
  int Mnpat = 0; long long Mpatlen = 0;
  // step1's LDS window of operand records: when every K pattern offset is
  // dz*vnxy + dy*vnx + dx with |dx|, |dy|, |dz| <= 2 for the velocity grid of
  // a structured mesh (vnx = 2 pg_nx - 1 nodes a line), a block of rows r0..
  // r0 + 31 reads its columns from 25 runs of 36 records staged in LDS, run
  // (dy, dz) holding ids r0 - 2 + dy*vnx + dz*vnxy + [0, 36); Mpatslot holds
  // each pattern entry's slot 2 + dx + 36 ((dz + 2) 5 + dy + 2), so that the
  // entry of row r sits at slot (r - r0) + Mpatslot[..]. Null otherwise.
  int* Mpatslot = nullptr;
  int vnx = 0; long long vnxy = 0;
  bool use_step1_window = false;   // fsx_set_step1_window (measured slower at 64^3)
  int* Grow = nullptr; int* Gcol = nullptr;
  double* G = nullptr;           // nnzG*3, (G1, G2, G3) interleaved per nonzero
  int* GTrow = nullptr; int* GTcol = nullptr; int* GTpos = nullptr;
  double* GT = nullptr;          // nnzG*3, transposed pattern
  int* Zrow = nullptr; int* Zcol = nullptr;
  double* Z = nullptr; double* Zdinv = nullptr;
  float* Zdinvf = nullptr;
  int zrow_max = 0;
  // structured pressure grid (x fastest), detected at fsx_upload_mesh; 0 when
  // the pressure numbering is not a grid (e.g. a permuted mesh)
  int pg_nx = 0, pg_ny = 0, pg_nz = 0;
  // the FP32 row classes of Z staged through LDS tiles of that grid (null
  // when the grid is not structured or Z is not a radius-2 stencil on it)
  fsx::TilePlan tile{};
  void* tile_mem[3] = {nullptr, nullptr, nullptr};
  double* tile_part = nullptr;   // 5 slots of tile.npart partial sums
  // Z without its structural zeros (entries below kPruneRel of the row's
  // largest magnitude), built by fsx_build_Z for the solvers:
  //  * FP64 rows Zr_row / Zr_val, columns through a pattern dictionary
  //    (Zr_pat, Zr_off as above) or, when the rows do not compress, Zr_col;
  //  * the FP32 copy as row classes when its rows take few distinct lists of
  //    (column offset, FP32 value): a class's entries sit in Zc_ent as
  //    (offset, float bits); null otherwise (the FP32 solvers then read
  //    Zr_valf on the rows of the FP64 form).
  int* Zr_row = nullptr; double* Zr_val = nullptr; int* Zr_col = nullptr;
  float* Zr_valf = nullptr;   // Zr_val rounded to FP32: the FP32 solvers' Z where the rows form no classes
  int* Zr_pat = nullptr; int* Zr_off = nullptr;
  long long nnzZr = 0; int Zr_npat = 0; long long Zr_patlen = 0;
  int2* Zc_ent = nullptr;
  int Zc_n = 0; long long Zc_len = 0;
  // the class rows' launch plan: the rows grouped by class (ascending within
  // a class) in Zc_rows, cut into chunks of at most 64 rows of one class:
  // (first index into Zc_rows, row count, class entry start, entry count)
  int4* Zc_chunk = nullptr; int Zc_nchunk = 0;
  int* Zc_rows = nullptr;
  // FP64 Z on the classes: each row's values minus their FP32 rounding, as
  // FP32 (Z64 = class value + delta to within 2^-48 of the value), chunk c's
  // block at Zc_doff[c] laid out entry-major, 64 lanes a row (the FP64
  // residuals of the mixed and learned solvers read these instead of Zr_val)
  float* Zc_delta = nullptr; long long* Zc_doff = nullptr;
  // lumped mass and BCs (device, NN-long; the reference replicates them x3)
  double* MdInv = nullptr; double* MdOrigInv = nullptr; double* Md = nullptr;
  unsigned char* bcflag = nullptr;
  int nbc = 0; int* bcnodes = nullptr; double* bcvals = nullptr;
  int zeroP = -1;
  // state (device). Velocity-sized vectors (3NN) are node-major: u, v, w of
  // node i at 3i, 3i+1, 3i+2, so the gathers of the K, G and G^T row kernels
  // and of the element kernels fetch one 24-byte run per node instead of
  // three scattered words. fsx_download / fsx_upload_vec convert to and from
  // the reference's [u|v|w] layout.
  double *Un = nullptr, *Unp1 = nullptr, *Unp1_prev = nullptr, *UnpHalf = nullptr, *UnpHalf_prev = nullptr;
  double *Acc = nullptr, *Acc_prev = nullptr, *KAcc = nullptr, *R1 = nullptr;
  double *Pn = nullptr, *Pnp1 = nullptr, *Pnp1_prev = nullptr, *Pdot = nullptr, *R2 = nullptr;
  // solver work (device)
  double *wr = nullptr, *wz = nullptr, *wp = nullptr, *wq = nullptr, *wh = nullptr;    // FP64 PCG (+ FP64 residual)
  double *w0r = nullptr, *w0e = nullptr;   // warm start: b - Z x0 and the correction
  // FSX_X0_PROJECTION: the last kProjMax solutions x_j and their products
  // Z x_j (slot i at proj_x + i NNp, proj_b + i NNp), the window's Gram
  // entries G[i][j] = x_i . Z x_j on the host, and the slots whose row and
  // column of G are still to be formed
  double *proj_x = nullptr, *proj_b = nullptr, *proj_part = nullptr;
  int proj_n = 0, proj_next = 0, proj_pending = 0;   // pending: bit mask of slots
  std::vector<double> proj_G;
  // the FP64 residual b - Z x0 the projection start formed beside x0 (the
  // buffer the next warm start would compute it into), or null
  double* warm_r = nullptr;
  int x0_policy = FSX_X0_ZERO;
  float *fr = nullptr, *fz = nullptr, *fp = nullptr, *fq = nullptr, *fx = nullptr;     // FP32 PCG
  double* fpq = nullptr; int fpq_n = 0;   // learned solver: partial sums of p.Zp from its SpMV
  double* partials = nullptr;    // 8 * kMaxGrid, per-block partial sums of the reduction kernels
  fsx::SolveScalars* sc = nullptr;
  fsx::StepScalars* st = nullptr;
  fsx::Published* h_pub = nullptr;     // host-coherent, written by k_publish
  fsx::Published* d_pub = nullptr;     // its device address
  unsigned pub_seq = 0;
  fsx::SolveScalars* h_sc = nullptr;   // pinned host mirrors
  fsx::StepScalars* h_st = nullptr;
  double* h_mon = nullptr;
  double* d_mon = nullptr;
  fsx::ShapeTables tab{};
  int conv_grid = 0;             // blocks of the persistent convection kernel (0: not yet queried)
  int conv_grid_f = 0;           // the same for the FP32 convection
  double* tab_mem = nullptr;
  fsx::ShapeTablesF tabf{};     // FP32 copy of tab
  float* tabf_mem = nullptr;
  // accounting
  std::vector<void*> allocs;
  size_t bytes = 0;
  bool timing = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  fsx_timing tm{};
  int last_iters = 0;
  int last_updates = 0;
  double last_relres = 0.0;
  // learned preconditioner
  struct fsx_unet* unet = nullptr;
  // multifrontal direct solve in a nested-dissection order (fsx_sparse.hip)
  fsx::MfState* mf = nullptr;
  int direct_order = FSX_DIRECT_ND;
  int learned_maxit = 400;       // flexible-CG iterations before the Jacobi fallback
  // preconditioner form (fsx_unet_form): 1 the two-level multiplicative form
  // on the LDS tiles (default where the tile plan exists), 0 the additive form
  int learned_form = 1;
  float learned_omega = 0.8f;    // Jacobi smoothing weight of the two-level form
  float* fr2 = nullptr;          // the tiled flexible CG's second residual and direction buffers,
  float* fp2 = nullptr;          // the network's input r1 and output z2 (NNp each)
  float* fr1 = nullptr;
  float* fz2 = nullptr;
  // flexible iterations of each segment (between reliable updates) of the
  // last tiled learned solve: the next solve enqueues that many before it
  // reads the scalars back
  int learned_seg[16] = {0};
  int last_learned_iters = 0;
  int last_fallback = 0;
  void (*capture_hook)(fsx_ctx*, const double* b, const double* x) = nullptr;
  // the scratch array of fsx_solve_pressure_noalloc on this device (its
  // device address, from hipGetSymbolAddress at fsx_create)
  void* scratch = nullptr;
  // iteration graphs: FP64 PCG, mixed inner PCG, learned flexible PCG
  bool use_graphs = false;      // measured on par with individual launches on this path
  fsx::IterGraph graph_fp64, graph_mixed, graph_learned;
};

// error helpers (host)
#define FSX_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t _e = (call);                                                             \
    if (_e != hipSuccess) {                                                             \
      ctx->err = std::string(#call) + " -> " + hipGetErrorString(_e);                  \
      return FSX_EHIP;                                                                  \
    }                                                                                   \
  } while (0)

#define FSX_TRY(call)                 \
  do {                                \
    int _rc = (call);                 \
    if (_rc != FSX_OK) return _rc;    \
  } while (0)

namespace fsx {
int dev_alloc(fsx_ctx* ctx, void** p, size_t bytes);
template <class T>
int alloc(fsx_ctx* ctx, T** p, size_t n) {
  return dev_alloc(ctx, reinterpret_cast<void**>(p), n * sizeof(T));
}
void dev_free(fsx_ctx* ctx, void* p);
int fail(fsx_ctx* ctx, int code, const std::string& msg);
int launch_check(fsx_ctx* ctx, const char* what);
int row_grid(long long nrows, int rows_per_block);

// kernel entry points implemented in the .hip files (host-side launchers)
int launch_assemble_static(fsx_ctx* ctx);
int launch_lump(fsx_ctx* ctx);
int launch_build_Z(fsx_ctx* ctx);
int launch_z_refresh(fsx_ctx* ctx);
int launch_convection(fsx_ctx* ctx);
int launch_step1(fsx_ctx* ctx, int iter);
int records_begin_step(fsx_ctx* ctx);
int ensure_fpq(fsx_ctx* ctx);
int direct_factor(fsx_ctx* ctx);
int direct_apply(fsx_ctx* ctx, const double* b, double* x);
inline void drop_records(fsx_ctx* ctx) {
  ctx->urec_ok = ctx->urec_prev_ok = ctx->arec_ok = ctx->arec_prev_ok = ctx->drec_ok = ctx->ucmp_ok = false;
  ctx->ucmp_un = false;
}
int launch_step2_rhs(fsx_ctx* ctx, int iter);
int launch_pnp1(fsx_ctx* ctx);
int launch_step3(fsx_ctx* ctx, int iter);
int launch_kacc(fsx_ctx* ctx);
int settle_deferred(fsx_ctx* ctx);
int launch_convection_now(fsx_ctx* ctx);
int elem_tables_setup(fsx_ctx* ctx, const double* tables, int NGP);
bool step1_elements_ok(const fsx_ctx* ctx);
int launch_step1_elements(fsx_ctx* ctx, bool kacc, bool gload);
int launch_elem_pass(fsx_ctx* ctx, bool kacc);
int mf_factor(fsx_ctx* ctx);
int mf_apply(fsx_ctx* ctx, const double* b, double* x);
void mf_free(fsx_ctx* ctx);
void mf_invalidate(fsx_ctx* ctx);
int mf_info(const fsx_ctx* ctx, double* l_bytes, double* flops, int* fronts, int* depth, double* factor_ms);
int launch_nodes_pass(fsx_ctx* ctx, bool kacc, bool gload);
int refresh_matrix_copies(fsx_ctx* ctx);
// empties the FSX_X0_PROJECTION window (fsx_solve.hip)
void proj_reset(fsx_ctx* ctx);
// host spin until the pinned publication carries sequence number seq
int wait_published(fsx_ctx* ctx, unsigned seq, const char* what);
// whether step3 and step2's right-hand side read G and G^T as row classes
inline bool g_classes_on(const fsx_ctx* ctx) { return ctx->Gf && ctx->GTf && ctx->Gc_row && ctx->use_g_classes; }
int launch_mass_fac(fsx_ctx* ctx, const int* P8, double* fac);
int elements_setup(fsx_ctx* ctx);
int mass_update_step1(fsx_ctx* ctx);
int mass_update_step3(fsx_ctx* ctx);
int launch_maxacc(fsx_ctx* ctx);
int launch_fill(fsx_ctx* ctx, double* p, double v, long long n);
int launch_copy(fsx_ctx* ctx, double* dst, const double* src, long long n);
int launch_copy2(fsx_ctx* ctx, double* d1, const double* s1, long long n1, double* d2, const double* s2, long long n2);
int pressure_solve(fsx_ctx* ctx, const double* b, double* x, int* iters, double* relres);
int spmv_host_op(fsx_ctx* ctx, int which, const double* dx, double* dy, double alpha, double beta);
int launch_add(fsx_ctx* ctx, double* x, const double* d, int n);
int launch_residual_only(fsx_ctx* ctx, double* x, const double* b, int* stop, double* rr_hi);
int build_tile_plan(fsx_ctx* ctx, const std::vector<int>& cls, const std::vector<int>& cp, const std::vector<int>& co,
                    const std::vector<unsigned>& cv);
struct ZTensor;
// the tile plan of the Kronecker form (fsx_stencil.hip); FSX_OK with no
// plan (ntiles 0) when a tile does not fit LDS
int build_tile_plan_tensor(fsx_ctx* ctx, const ZTensor& zt);
int launch_tile(fsx_ctx* ctx, int mode, const TileVecs& v);
int launch_warm_residual(fsx_ctx* ctx, const double* x, const double* b, double* rh);
int launch_tile_rescan(fsx_ctx* ctx, const double* rh, float* r, const float* p);
void unet_free(fsx_ctx* ctx);
// *ctx->h_sc <- the solve scalars (sc, or the context's own when null) after
// all work enqueued so far
int read_scalars(fsx_ctx* ctx, const SolveScalars* sc = nullptr);
// the device address of the allocation-free entry point's scratch array on
// the context's device (once per device, at fsx_create)
int scratch_setup(fsx_ctx* ctx);
void iter_graph_reset(IterGraph& g);

// Enqueue iterations j .. j + count - 1 (rounded up to whole graphs when
// graphs are on; iterations after the stop flag is set are no-ops).
// issue(jj) enqueues one iteration on ctx->stream and returns a status.
// *enqueued returns how many iterations were enqueued (the next start).
template <class F>
int run_iterations(fsx_ctx* ctx, IterGraph& G, unsigned long long key, int j, int count, int* enqueued, F&& issue) {
  if (!ctx->use_graphs) {
    for (int c = 0; c < count; c++) FSX_TRY(issue(j + c));
    *enqueued = count;
    return launch_check(ctx, "solver iterations");
  }
  if (G.key != key) { iter_graph_reset(G); G.key = key; }
  const int par = j & 1;
  if (!G.exec[par]) {
    FSX_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int rc = FSX_OK;
    for (int i = 0; i < kGraphBlock && rc == FSX_OK; i++) rc = issue(par + i);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc != FSX_OK) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return fail(ctx, FSX_EHIP, std::string("hipStreamEndCapture -> ") + hipGetErrorString(e));
    G.graph[par] = g;
    FSX_HIP(hipGraphInstantiate(&G.exec[par], g, nullptr, nullptr, 0));
  }
  int c = 0;
  for (; c < count; c += kGraphBlock) FSX_HIP(hipGraphLaunch(G.exec[par], ctx->stream));
  *enqueued = c;
  return FSX_OK;
}
}  // namespace fsx
