// fsx_solve.hip -- CSR row kernels of the time step (fused SpMV + elementwise
// epilogues), reductions and the Jacobi-PCG pressure solvers.
//
// Row kernels: one block per 256/LPR consecutive rows, a group of LPR lanes
// per row. Each lane walks its row in rounds of U entries whose column and
// value loads are all issued before the dependent gathers and FMAs, so a
// wavefront keeps several HBM requests in flight. (Measured on the 64^3 Z,
// 32.5 M nonzeros: this form streams at 5.2 TB/s; grid-stride row chunks,
// non-temporal loads or a reduction fused into the SpMV were each slower.)
//
// Reductions never use fences or atomics: a fixed grid of kRedGrid blocks
// writes one partial per block and the next kernel sums them redundantly, in
// a fixed order, in every block (fsx_device.h). Results are bitwise
// reproducible.
#include "fsx_device.h"
#include "fsx_gclass.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace fsx {

// ------------------------------------------------------------ elementwise
__global__ void k_fill(double* __restrict__ p, double v, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_copy(double* __restrict__ d, const double* __restrict__ s, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) d[i] = s[i];
}

// two copies in one launch: d1 = s1 (n1 values), then d2 = s2 (n2 values)
__global__ void k_copy2(double* __restrict__ d1, const double* __restrict__ s1, long long n1, double* __restrict__ d2,
                        const double* __restrict__ s2, long long n2) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n1 + n2; i += (long long)gridDim.x * blockDim.x) {
    if (i < n1) d1[i] = s1[i];
    else d2[i - n1] = s2[i - n1];
  }
}

static int ew_grid(long long n) {
  long long g = (n + kBlock - 1) / kBlock;
  return (int)std::max(1LL, std::min(g, 4096LL));
}

static unsigned rows_grid(long long nrows, int lpr) {
  const long long rpb = kBlock / lpr;
  return (unsigned)std::max(1LL, (nrows + rpb - 1) / rpb);
}

int launch_fill(fsx_ctx* ctx, double* p, double v, long long n) {
  hipLaunchKernelGGL(k_fill, dim3(ew_grid(n)), dim3(kBlock), 0, ctx->stream, p, v, n);
  return launch_check(ctx, "k_fill");
}
int launch_copy(fsx_ctx* ctx, double* dst, const double* src, long long n) {
  hipLaunchKernelGGL(k_copy, dim3(ew_grid(n)), dim3(kBlock), 0, ctx->stream, dst, src, n);
  return launch_check(ctx, "k_copy");
}
int launch_copy2(fsx_ctx* ctx, double* d1, const double* s1, long long n1, double* d2, const double* s2, long long n2) {
  hipLaunchKernelGGL(k_copy2, dim3(ew_grid(n1 + n2)), dim3(kBlock), 0, ctx->stream, d1, s1, n1, d2, s2, n2);
  return launch_check(ctx, "k_copy2");
}

// Row of this lane group and the lane inside the group. Blocks take rows in
// launch order: workgroups are dealt to the 8 XCDs round robin, so all XCDs
// sweep the same region of rows at once and share the x-gathers through the
// Infinity Cache. (Handing each XCD its own contiguous slab of rows instead
// measured slower on the 64^3 cavity: K*Acc 0.67 -> 1.00 ms, step3 0.32 ->
// 0.48 ms.)
template <int LPR>
__device__ __forceinline__ long long my_row(int* lane) {
  *lane = threadIdx.x % LPR;
  return (long long)blockIdx.x * (kBlock / LPR) + threadIdx.x / LPR;
}

// ------------------------------------------------------------ step 1
// R1 <- (R1 - K UnpHalf_prev) - G Pn; BC rows 0; UnpHalf = Un + dt R1 MdInv.
// (BCH:3746-3783; the convection part of R1 is already in R1.) With KACC the
// same pass over K also forms KAcc = K Acc_prev (BCH:2981-3035), which the
// reference computes at the end of the previous inner iteration: K's columns
// and values are read once for six right-hand sides instead of twice for
// three. Pn is fixed over the inner iterations of a time step, so G Pn is the
// same in each: GMODE 1 forms it and keeps it in GPn, GMODE 2 reads it back
// (24 bytes per row instead of G's 28 per nonzero), GMODE 0 only forms it.
// With REC the K operands are gathered from the FP32 shadows Urp (and Arp)
// of UnpHalf_prev (and Acc_prev), one 16-byte load per column and vector
// instead of 24 bytes, the two of a column side by side in one 32-byte
// record (node i at 8i floats); the products and sums stay FP64. A non-null Uout
// receives UnpHalf rounded to FP32, the shadow the next inner iteration reads,
// and a non-null Dout step2's operand UnpHalf / dt^2 (- MdOrigInv KAcc with
// KACC) rounded to FP32.
// With WIN (REC and KPAT only) the block first stages the operand records of
// its column window in LDS (25 runs of 36 nodes, see Mpatslot) and the rows
// read them there; the products and their order are the same.
template <int LPR, int UK, int UG, bool KACC, bool KPAT, int GMODE, class VT, bool REC, bool WIN = false>
__global__ __launch_bounds__(kBlock) void k_step1(int NN, const int* __restrict__ Mrow, ColSrc mcs,
                                                  const VT* __restrict__ K, const int* __restrict__ Grow,
                                                  const int* __restrict__ Gcol, const VT* __restrict__ G,
                                                  const double* __restrict__ Uprev, const double* __restrict__ Pn,
                                                  const double* __restrict__ Un, double* __restrict__ R1,
                                                  const double* __restrict__ MdInv, const unsigned char* __restrict__ bc,
                                                  double* __restrict__ UnpHalf, double dt,
                                                  const double* __restrict__ Accp, double* __restrict__ KAcc,
                                                  double* __restrict__ GPn, const float4* __restrict__ Urp,
                                                  const float4* __restrict__ Arp, float* __restrict__ Uout,
                                                  const double* __restrict__ MdOrigInv, double o2,
                                                  float* __restrict__ Dout, int us, const int* __restrict__ pslot,
                                                  int vnx, long long vnxy) {
  static_assert(!WIN || (REC && KPAT), "the window reads FP32 records through the K patterns");
  constexpr int kWinRows = kBlock / LPR, kWinLen = kWinRows + 4, kWinSlots = 25 * kWinLen;
  __shared__ float4 win[WIN ? 2 * kWinSlots : 1];
  int l;
  const long long r = my_row<LPR>(&l);
  const long long r0 = (long long)blockIdx.x * kWinRows;
  if constexpr (WIN) {
    for (int t = threadIdx.x; t < kWinSlots; t += kBlock) {
      const int run = t / kWinLen, pos = t - run * kWinLen;
      const long long id = r0 - 2 + pos + (long long)(run % 5 - 2) * vnx + (long long)(run / 5 - 2) * vnxy;
      if (id >= 0 && id < NN) {
        win[(size_t)us * t] = Urp[(size_t)us * id];
        if (us == 2) win[2 * (size_t)t + 1] = Urp[2 * (size_t)id + 1];
      }
    }
    __syncthreads();
  }
  double k0 = 0, k1 = 0, k2 = 0, g0 = 0, g1 = 0, g2 = 0, a0 = 0, a1 = 0, a2 = 0;
  if (r < NN) {
    const int rs = Mrow[r], re = Mrow[r + 1];
    const int cb = col_base<KPAT>(mcs, r, rs);
    for (int base = rs + l; base < re; base += UK * LPR) {
      int c[UK];
      double a[UK];
#pragma unroll
      for (int j = 0; j < UK; j++) {
        const int k = base + j * LPR, kk = k < re ? k : re - 1;
        if constexpr (WIN) c[j] = (int)(r - r0) + pslot[cb + kk];   // the entry's window slot
        else c[j] = col_at<KPAT>(mcs, r, cb, kk);
        a[j] = k < re ? (double)K[kk] : 0.0;
      }
#pragma unroll
      for (int j = 0; j < UK; j++) {
        if constexpr (WIN) {
          const float4 u = win[(size_t)us * c[j]];
          k0 += a[j] * (double)u.x; k1 += a[j] * (double)u.y; k2 += a[j] * (double)u.z;
          if constexpr (KACC) {
            const float4 v = win[2 * (size_t)c[j] + 1];
            a0 += a[j] * (double)v.x; a1 += a[j] * (double)v.y; a2 += a[j] * (double)v.z;
          }
        } else if constexpr (REC) {
          const float4 u = Urp[(size_t)us * c[j]];
          k0 += a[j] * (double)u.x; k1 += a[j] * (double)u.y; k2 += a[j] * (double)u.z;
          if constexpr (KACC) {
            const float4 v = Arp[2 * (size_t)c[j]];
            a0 += a[j] * (double)v.x; a1 += a[j] * (double)v.y; a2 += a[j] * (double)v.z;
          }
        } else {
          const size_t x = 3 * (size_t)c[j];
          k0 += a[j] * Uprev[x]; k1 += a[j] * Uprev[x + 1]; k2 += a[j] * Uprev[x + 2];
          if constexpr (KACC) { a0 += a[j] * Accp[x]; a1 += a[j] * Accp[x + 1]; a2 += a[j] * Accp[x + 2]; }
        }
      }
    }
    if constexpr (GMODE != 2) {
      const int gs = Grow[r], ge = Grow[r + 1];
      for (int base = gs + l; base < ge; base += UG * LPR) {
        int c[UG];
        double b0[UG], b1[UG], b2[UG];
#pragma unroll
        for (int j = 0; j < UG; j++) {
          const int k = base + j * LPR, kk = k < ge ? k : ge - 1;
          c[j] = Gcol[kk];
          const bool ok = k < ge;
          b0[j] = ok ? (double)G[3 * (size_t)kk] : 0.0;
          b1[j] = ok ? (double)G[3 * (size_t)kk + 1] : 0.0;
          b2[j] = ok ? (double)G[3 * (size_t)kk + 2] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < UG; j++) {
          const double p = Pn[c[j]];
          g0 += b0[j] * p; g1 += b1[j] * p; g2 += b2[j] * p;
        }
      }
    }
  }
  k0 = group_sum<LPR>(k0); k1 = group_sum<LPR>(k1); k2 = group_sum<LPR>(k2);
  if constexpr (GMODE != 2) { g0 = group_sum<LPR>(g0); g1 = group_sum<LPR>(g1); g2 = group_sum<LPR>(g2); }
  if constexpr (KACC) { a0 = group_sum<LPR>(a0); a1 = group_sum<LPR>(a1); a2 = group_sum<LPR>(a2); }
  if (r < NN && l < 3) {
    const double kk = l == 0 ? k0 : (l == 1 ? k1 : k2);
    const size_t i = 3 * (size_t)r + l;
    double gg;
    if constexpr (GMODE == 2) {
      gg = GPn[i];
    } else {
      gg = l == 0 ? g0 : (l == 1 ? g1 : g2);
      if constexpr (GMODE == 1) GPn[i] = gg;
    }
    double R = (R1[i] - kk) - gg;
    if (bc[r]) R = 0.0;
    R1[i] = R;
    const double uh = Un[i] + dt * R * MdInv[r];
    UnpHalf[i] = uh;
    if (Uout) Uout[8 * (size_t)r + l] = (float)uh;
    double ka = 0.0;
    if constexpr (KACC) { ka = l == 0 ? a0 : (l == 1 ? a1 : a2); KAcc[i] = ka; }
    if (Dout) {   // step2's operand, as k_step2_rhs forms it
      double d = o2 * uh;
      if constexpr (KACC) d = d - MdOrigInv[r] * ka;
      Dout[4 * (size_t)r + l] = (float)d;
    }
  }
}

int launch_step1(fsx_ctx* ctx, int iter) {
  (void)iter;
  if (ctx->conv_pending) {
    // the convection recorded by fsx_convection: formed here with the K
    // products element by element, or on its own when that form no longer applies
    if (step1_elements_ok(ctx)) {
      const bool kacc = ctx->kacc_pending;
      FSX_TRY(launch_step1_elements(ctx, kacc, ctx->gpn_valid));
      ctx->conv_pending = ctx->kacc_pending = false;
      ctx->gpn_valid = true;
      ctx->urec_ok = ctx->drec_ok = true;
      ctx->drec_kacc = kacc;
      return FSX_OK;
    }
    FSX_TRY(launch_convection_now(ctx));
  }
  // 8 lanes x 2 K entries in flight per row. Fused with K*Acc at 64^3
  // (node-major vectors): 8x2 1.064 ms, 16x2 1.057, 8x4 1.077, 16x4 1.092,
  // 8x8 1.150, 8x16 1.217, 16x8 1.224, 4x16 1.352, 32x2 1.287
  constexpr int LPR = 8, UK = 2;
  const dim3 grid(rows_grid(ctx->NN, LPR));
  const ColSrc mcs{ctx->Mcol, ctx->Mrowpat, ctx->Mpatoff};
  const bool kacc = ctx->kacc_pending;
  ctx->kacc_pending = false;
  // G Pn: formed and kept by the first step1 after Pn changed, read back after
  const bool gload = ctx->gpn_valid;
  const bool f32 = ctx->Kf != nullptr;   // FP32 matrix values (fsx_set_matrix_precision)
  // FP32 operand shadows: written with lumped mass (the consistent update
  // rewrites UnpHalf afterwards), read when they hold the prev vectors
  const bool wrec = f32 && ctx->rec_mem && ctx->mass_mode == FSX_MASS_LUMPED;
  const bool rrec = wrec && ctx->urec_prev_ok && (!kacc || ctx->arec_prev_ok);
  // the first inner iteration (no K Acc) reads the compact shadow of Un:
  // the interleaved records would bring Acc's unused half along
  const bool cmp = rrec && !kacc && ctx->ucmp_ok;
  const float4* urp = reinterpret_cast<const float4*>(cmp ? ctx->Ucmp : ctx->Urec_prev);
  const int us = cmp ? 1 : 2;   // float4 records a node
  const float4* arp = reinterpret_cast<const float4*>(ctx->Arec_prev);
  float* uout = wrec ? ctx->Urec : nullptr;
  float* dout = wrec ? ctx->Drec : nullptr;
  const double o2 = 1.0000000000000000 / (ctx->dt * ctx->dt);
  // the LDS window: the K patterns decompose on the velocity grid, and the
  // Acc records sit beside the UnpHalf ones (interleaved shadows)
  const bool win = rrec && mcs.row_pat && ctx->Mpatslot && ctx->use_step1_window &&
                   (!kacc || (us == 2 && arp == urp + 1));
#define STEP1_LAUNCH_T(KA, KP, GM, VT, RC, KV, GV, WN)                                                                \
  hipLaunchKernelGGL((k_step1<LPR, UK, 2, KA, KP, GM, VT, RC, WN>), grid, dim3(kBlock), 0, ctx->stream, ctx->NN,       \
                     ctx->Mrow, mcs, KV, ctx->Grow, ctx->Gcol, GV, ctx->UnpHalf_prev, ctx->Pn, ctx->Un, ctx->R1,        \
                     ctx->MdInv, ctx->bcflag, ctx->UnpHalf, ctx->dt, KA ? ctx->Acc_prev : nullptr,                     \
                     KA ? ctx->KAcc : nullptr, ctx->GPn, urp, arp, uout, ctx->MdOrigInv, o2, dout, us, ctx->Mpatslot,  \
                     ctx->vnx, ctx->vnxy)
#define STEP1_LAUNCH(KA, KP, GM) \
  do { if (rrec && KP && win) STEP1_LAUNCH_T(KA, KP, GM, float, true, ctx->Kf, ctx->Gf, KP); \
       else if (rrec) STEP1_LAUNCH_T(KA, KP, GM, float, true, ctx->Kf, ctx->Gf, false); \
       else if (f32) STEP1_LAUNCH_T(KA, KP, GM, float, false, ctx->Kf, ctx->Gf, false); \
       else STEP1_LAUNCH_T(KA, KP, GM, double, false, ctx->K, ctx->G, false); } while (0)
#define STEP1_G(KA, KP) \
  do { if (gload) STEP1_LAUNCH(KA, KP, 2); else STEP1_LAUNCH(KA, KP, 1); } while (0)
  if (mcs.row_pat) {
    if (kacc) STEP1_G(true, true); else STEP1_G(false, true);
  } else {
    if (kacc) STEP1_G(true, false); else STEP1_G(false, false);
  }
#undef STEP1_G
#undef STEP1_LAUNCH
#undef STEP1_LAUNCH_T
  FSX_TRY(launch_check(ctx, "k_step1"));
  ctx->gpn_valid = true;
  ctx->urec_ok = wrec;
  ctx->drec_ok = wrec;
  ctx->drec_kacc = kacc;
  if (ctx->mass_mode != FSX_MASS_LUMPED) return mass_update_step1(ctx);
  return FSX_OK;
}

// ------------------------------------------------------------ matrix precision
__global__ void k_d2f(long long n, const double* __restrict__ a, float* __restrict__ b) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

// G and G^T as row classes (fsx_gclass.h) from the FP64 assembly, when the
// matrix precision is 32 and the rows compress. A component below 2^-24 of
// its row's largest magnitude is dropped: on congruent elements such
// components are the rounding residue of contributions that cancel, and
// their differences alone would keep rows that are otherwise equal apart.
// The drop is below FP32's resolution of the row's largest term.
constexpr double kGClassPrune = 5.9604644775390625e-08;   // 2^-24
constexpr int kGClassMax = 65535;
constexpr long long kGClassMaxEnt = 1LL << 22;   // class entry starts fit 23 bits
constexpr int kGClassMaxLen = 255;               // row lengths fit 8 bits

static int build_g_classes(fsx_ctx* ctx) {
  // built once per assembly of G, when first wanted; kept while the
  // precision or the switch changes (g_classes_on decides their use)
  if (!ctx->use_g_classes || !ctx->Gf || ctx->gc_built) return FSX_OK;
  for (void* p : {(void*)ctx->Gc_row, (void*)ctx->Gc_ent, (void*)ctx->GTc_info, (void*)ctx->GTc_col, (void*)ctx->GTc_base,
                  (void*)ctx->GTc_val})
    if (p) dev_free(ctx, p);
  ctx->Gc_row = nullptr; ctx->Gc_ent = nullptr; ctx->GTc_info = nullptr; ctx->GTc_col = nullptr; ctx->GTc_base = nullptr;
  ctx->GTc_val = nullptr;
  ctx->gc_ncls = ctx->gtc_ncls = 0;
  ctx->gc_nent = ctx->gtc_nent = 0;
  ctx->gc_built = true;
  const int nn = ctx->NN, np = ctx->NNp;
  const long long nnz = ctx->nnzG;
  std::vector<int> rp(nn + 1), ci(nnz), trp(np + 1);
  std::vector<double> g(3 * (size_t)nnz);
  FSX_HIP(hipStreamSynchronize(ctx->stream));
  FSX_HIP(hipMemcpy(rp.data(), ctx->Grow, (size_t)(nn + 1) * sizeof(int), hipMemcpyDeviceToHost));
  FSX_HIP(hipMemcpy(ci.data(), ctx->Gcol, (size_t)nnz * sizeof(int), hipMemcpyDeviceToHost));
  FSX_HIP(hipMemcpy(g.data(), ctx->G, 3 * (size_t)nnz * sizeof(double), hipMemcpyDeviceToHost));
  FSX_HIP(hipMemcpy(trp.data(), ctx->GTrow, (size_t)(np + 1) * sizeof(int), hipMemcpyDeviceToHost));
  GClassForm F;
  // the form pays where the class lists are a small part of the operator:
  // at most an eighth of its entries
  const int rc = g_class_form(nn, np, rp.data(), ci.data(), g.data(), kGClassPrune, kGClassMax,
                              std::min(kGClassMaxEnt, nnz / 8), kGClassMaxLen, &F);
  if (rc < 0) return fail(ctx, FSX_EINVAL, "G's pattern is inconsistent");
  if (rc > 0 || F.trow_ptr != trp) return FSX_OK;   // no compression: the step kernels read Gf / GTf
  std::vector<int2> grow(nn);
  for (int r = 0; r < nn; r++)
    grow[r] = make_int2(F.row_base[r], (F.cls_ptr[F.row_cls[r]] << 8) | (rp[r + 1] - rp[r]));
  const size_t ne = F.ent_off.size();
  std::vector<float4> gent(ne);
  for (size_t t = 0; t < ne; t++) {
    float off;
    std::memcpy(&off, &F.ent_off[t], sizeof off);
    gent[t] = make_float4(off, F.ent_val[3 * t], F.ent_val[3 * t + 1], F.ent_val[3 * t + 2]);
  }
  std::vector<int> tinfo(np), tbase(F.toffsets ? np : 0);
  for (int p = 0; p < np; p++) {
    tinfo[p] = (F.tcls_ptr[F.trow_cls[p]] << 8) | (trp[p + 1] - trp[p]);
    if (F.toffsets) tbase[p] = F.tcol[trp[p]];
  }
  const size_t nte = F.tent_val.size() / 3;
  std::vector<float4> tval(nte);
  for (size_t t = 0; t < nte; t++) {
    float off = 0.0f;
    if (F.toffsets) std::memcpy(&off, &F.tent_off[t], sizeof off);
    tval[t] = make_float4(F.tent_val[3 * t], F.tent_val[3 * t + 1], F.tent_val[3 * t + 2], off);
  }
  FSX_TRY(alloc(ctx, &ctx->Gc_row, (size_t)nn));
  FSX_TRY(alloc(ctx, &ctx->Gc_ent, std::max<size_t>(1, ne)));
  FSX_TRY(alloc(ctx, &ctx->GTc_info, (size_t)np));
  FSX_TRY(alloc(ctx, &ctx->GTc_val, std::max<size_t>(1, nte)));
  FSX_HIP(hipMemcpy(ctx->Gc_row, grow.data(), (size_t)nn * sizeof(int2), hipMemcpyHostToDevice));
  if (ne) FSX_HIP(hipMemcpy(ctx->Gc_ent, gent.data(), ne * sizeof(float4), hipMemcpyHostToDevice));
  FSX_HIP(hipMemcpy(ctx->GTc_info, tinfo.data(), (size_t)np * sizeof(int), hipMemcpyHostToDevice));
  if (nte) FSX_HIP(hipMemcpy(ctx->GTc_val, tval.data(), nte * sizeof(float4), hipMemcpyHostToDevice));
  if (F.toffsets) {
    FSX_TRY(alloc(ctx, &ctx->GTc_base, (size_t)np));
    FSX_HIP(hipMemcpy(ctx->GTc_base, tbase.data(), (size_t)np * sizeof(int), hipMemcpyHostToDevice));
  } else {
    FSX_TRY(alloc(ctx, &ctx->GTc_col, std::max<size_t>(1, (size_t)nnz)));
    if (nnz) FSX_HIP(hipMemcpy(ctx->GTc_col, F.tcol.data(), (size_t)nnz * sizeof(int), hipMemcpyHostToDevice));
  }
  ctx->gc_ncls = (int)F.cls_ptr.size() - 1;
  ctx->gc_nent = (long long)ne;
  ctx->gtc_ncls = (int)F.tcls_ptr.size() - 1;
  ctx->gtc_nent = (long long)nte;
  return FSX_OK;
}

// The step kernels' FP32 copies of K, G and G^T (rounded to nearest) at
// matrix precision 32, after assembly or a K upload, and G's row classes
// where they form; at precision 64 the kernels read the FP64 values.
int refresh_matrix_copies(fsx_ctx* ctx) {
  ctx->Kf = ctx->Gf = ctx->GTf = nullptr;
  drop_records(ctx);
  if (ctx->mat_bits != 32 || !ctx->assembled) return FSX_OK;
  if (!ctx->Kf_mem) {
    FSX_TRY(alloc(ctx, &ctx->Kf_mem, (size_t)ctx->nnzM));
    FSX_TRY(alloc(ctx, &ctx->Gf_mem, 3 * (size_t)ctx->nnzG));
    FSX_TRY(alloc(ctx, &ctx->GTf_mem, 3 * (size_t)ctx->nnzG));
  }
  if (!ctx->rec_mem) {
    const size_t n4 = 4 * (size_t)ctx->NN;
    FSX_TRY(alloc(ctx, &ctx->rec_mem, 6 * n4));
    // UnpHalf's and Acc's shadows interleaved, 8 floats a node (u v w - a b c -),
    // so that step1's two gathers of a column fall in one 32-byte run
    ctx->Urec = ctx->rec_mem; ctx->Arec = ctx->rec_mem + 4;
    ctx->Urec_prev = ctx->rec_mem + 2 * n4; ctx->Arec_prev = ctx->rec_mem + 2 * n4 + 4;
    ctx->Drec = ctx->rec_mem + 4 * n4;
    ctx->Ucmp = ctx->rec_mem + 5 * n4;
  }
  hipLaunchKernelGGL(k_d2f, dim3(4096), dim3(kBlock), 0, ctx->stream, ctx->nnzM, ctx->K, ctx->Kf_mem);
  hipLaunchKernelGGL(k_d2f, dim3(4096), dim3(kBlock), 0, ctx->stream, 3 * ctx->nnzG, ctx->G, ctx->Gf_mem);
  hipLaunchKernelGGL(k_d2f, dim3(4096), dim3(kBlock), 0, ctx->stream, 3 * ctx->nnzG, ctx->GT, ctx->GTf_mem);
  FSX_TRY(launch_check(ctx, "k_d2f"));
  ctx->Kf = ctx->Kf_mem; ctx->Gf = ctx->Gf_mem; ctx->GTf = ctx->GTf_mem;
  return build_g_classes(ctx);
}

// FP32 shadow of a node-major velocity vector (zeros for a null v): node i
// at rec[stride i] (stride 2 in the interleaved 8-float records, 1 compact)
__global__ void k_rec(long long nn, const double* __restrict__ v, float4* __restrict__ rec, int stride) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nn; i += (long long)gridDim.x * blockDim.x)
    rec[stride * i] = v ? make_float4((float)v[3 * i], (float)v[3 * i + 1], (float)v[3 * i + 2], 0.0f)
                        : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// The shadows of UnpHalf_prev = Un and Acc_prev = 0 that fsx_begin_step sets
// (only where step1 and step3 keep the shadows up to date)
int records_begin_step(fsx_ctx* ctx) {
  drop_records(ctx);
  if (!ctx->Kf || !ctx->rec_mem || ctx->mass_mode != FSX_MASS_LUMPED) return FSX_OK;
  hipLaunchKernelGGL(k_rec, dim3(ew_grid(ctx->NN)), dim3(kBlock), 0, ctx->stream, (long long)ctx->NN, ctx->Un,
                     reinterpret_cast<float4*>(ctx->Urec_prev), 2);
  hipLaunchKernelGGL(k_rec, dim3(ew_grid(ctx->NN)), dim3(kBlock), 0, ctx->stream, (long long)ctx->NN, nullptr,
                     reinterpret_cast<float4*>(ctx->Arec_prev), 2);
  hipLaunchKernelGGL(k_rec, dim3(ew_grid(ctx->NN)), dim3(kBlock), 0, ctx->stream, (long long)ctx->NN, ctx->Un,
                     reinterpret_cast<float4*>(ctx->Ucmp), 1);
  FSX_TRY(launch_check(ctx, "k_rec"));
  ctx->urec_prev_ok = ctx->arec_prev_ok = ctx->ucmp_ok = ctx->ucmp_un = true;
  return FSX_OK;
}

// ------------------------------------------------------------ K * Acc_prev
// KAcc_c = K Acc_prev_c, c = 0..2, one pass over K (BCH:2981-3035)
template <int LPR, int U, bool KPAT, class VT>
__global__ __launch_bounds__(kBlock) void k_k3(int NN, const int* __restrict__ Mrow, ColSrc mcs,
                                               const VT* __restrict__ K, const double* __restrict__ x,
                                               double* __restrict__ y) {
  int l;
  const long long r = my_row<LPR>(&l);
  double k0 = 0, k1 = 0, k2 = 0;
  if (r < NN) {
    const int rs = Mrow[r], re = Mrow[r + 1];
    const int cb = col_base<KPAT>(mcs, r, rs);
    for (int base = rs + l; base < re; base += U * LPR) {
      int c[U];
      double a[U];
#pragma unroll
      for (int j = 0; j < U; j++) {
        const int k = base + j * LPR, kk = k < re ? k : re - 1;
        c[j] = col_at<KPAT>(mcs, r, cb, kk);
        a[j] = k < re ? (double)K[kk] : 0.0;
      }
#pragma unroll
      for (int j = 0; j < U; j++) {
        const size_t xi = 3 * (size_t)c[j];
        k0 += a[j] * x[xi]; k1 += a[j] * x[xi + 1]; k2 += a[j] * x[xi + 2];
      }
    }
  }
  k0 = group_sum<LPR>(k0); k1 = group_sum<LPR>(k1); k2 = group_sum<LPR>(k2);
  if (r < NN && l < 3) y[3 * (size_t)r + l] = l == 0 ? k0 : (l == 1 ? k1 : k2);
}

int launch_kacc(fsx_ctx* ctx) {
  constexpr int LPR = 16;
  const ColSrc mcs{ctx->Mcol, ctx->Mrowpat, ctx->Mpatoff};
  const dim3 grid(rows_grid(ctx->NN, LPR));
  if (ctx->Kf) {
    if (mcs.row_pat)
      hipLaunchKernelGGL((k_k3<LPR, 4, true, float>), grid, dim3(kBlock), 0, ctx->stream, ctx->NN, ctx->Mrow, mcs, ctx->Kf,
                         ctx->Acc_prev, ctx->KAcc);
    else
      hipLaunchKernelGGL((k_k3<LPR, 4, false, float>), grid, dim3(kBlock), 0, ctx->stream, ctx->NN, ctx->Mrow, mcs, ctx->Kf,
                         ctx->Acc_prev, ctx->KAcc);
  } else {
    if (mcs.row_pat)
      hipLaunchKernelGGL((k_k3<LPR, 4, true, double>), grid, dim3(kBlock), 0, ctx->stream, ctx->NN, ctx->Mrow, mcs, ctx->K,
                         ctx->Acc_prev, ctx->KAcc);
    else
      hipLaunchKernelGGL((k_k3<LPR, 4, false, double>), grid, dim3(kBlock), 0, ctx->stream, ctx->NN, ctx->Mrow, mcs, ctx->K,
                         ctx->Acc_prev, ctx->KAcc);
  }
  ctx->kacc_pending = false;
  ctx->drec_ok = false;   // step2's FP32 operand holds the KAcc before this one
  return launch_check(ctx, "k_k3");
}

// KAcc = K Acc_prev is deferred from fsx_kacc to the next step1, which forms
// it in its own pass over K (or over the elements); every other entry point
// that could read KAcc or change Acc_prev first settles a pending product
// with k_k3. A convection recorded for the element form of step1 is formed
// the same way (into R1) before anything else reads R1 or changes the state.
int settle_deferred(fsx_ctx* ctx) {
  if (ctx->conv_pending) FSX_TRY(launch_convection_now(ctx));
  return ctx->kacc_pending ? launch_kacc(ctx) : FSX_OK;
}

// ------------------------------------------------------------ step 2 RHS
// R2[p] = sum_c G_c^T d_c, d_c = UnpHalf_c / dt^2 - [iter>1] MdOrigInv KAcc_c,
// summed as (c0 + c1) + c2 like the three csrmv calls of BCH:3864-3868;
// R2[p0] = 0 for p0 > 0 (BCH:4180-4184).
// With DREC the operand d is gathered as step1's FP32 shadow of it (16 bytes
// a column instead of 48 plus MdOrigInv); sums stay FP64.
template <int LPR, int U, class VT, bool DREC, int CLS = 0>
__global__ __launch_bounds__(kBlock) void k_step2_rhs(int NNp, int NN, const int* __restrict__ GTrow,
                                                      const int* __restrict__ GTcol, const VT* __restrict__ GT,
                                                      const double* __restrict__ UnpHalf, const double* __restrict__ MdOrigInv,
                                                      const double* __restrict__ KAcc, int use_kacc, double o2,
                                                      int zeroP, double* __restrict__ R2, const float4* __restrict__ Drec,
                                                      const int* __restrict__ GTc_info = nullptr,
                                                      const float4* __restrict__ GTc_val = nullptr) {
  int l;
  const long long p = my_row<LPR>(&l);
  double s0 = 0, s1 = 0, s2 = 0;
  if (p < NNp) {
    // class forms (CLS 1, 2): values from the row's class entries; the
    // columns from GTcol in class order (1), or the class entries' offsets
    // from the row's first column, GTcol[p] (2)
    int rs, re, cs = 0, c0 = 0;
    if constexpr (CLS == 2) {
      const int info = GTc_info[p];
      rs = 0; re = info & 255; cs = info >> 8; c0 = GTcol[p];
    } else {
      rs = GTrow[p]; re = GTrow[p + 1];
      if constexpr (CLS == 1) cs = (GTc_info[p] >> 8) - rs;
    }
    for (int base = rs + l; base < re; base += U * LPR) {
      int c[U];
      double a0[U], a1[U], a2[U];
#pragma unroll
      for (int j = 0; j < U; j++) {
        const int k = base + j * LPR, kk = k < re ? k : re - 1;
        const bool ok = k < re;
        if constexpr (CLS > 0) {
          const float4 v = GTc_val[cs + kk];
          c[j] = CLS == 2 ? c0 + __float_as_int(v.w) : GTcol[kk];
          a0[j] = ok ? (double)v.x : 0.0;
          a1[j] = ok ? (double)v.y : 0.0;
          a2[j] = ok ? (double)v.z : 0.0;
        } else {
          c[j] = GTcol[kk];
          a0[j] = ok ? (double)GT[3 * (size_t)kk] : 0.0;
          a1[j] = ok ? (double)GT[3 * (size_t)kk + 1] : 0.0;
          a2[j] = ok ? (double)GT[3 * (size_t)kk + 2] : 0.0;
        }
      }
#pragma unroll
      for (int j = 0; j < U; j++) {
        const int i = c[j];
        if constexpr (DREC) {
          const float4 d = Drec[i];
          s0 += a0[j] * (double)d.x; s1 += a1[j] * (double)d.y; s2 += a2[j] * (double)d.z;
        } else {
          const size_t x = 3 * (size_t)i;
          double d0 = o2 * UnpHalf[x], d1 = o2 * UnpHalf[x + 1], d2 = o2 * UnpHalf[x + 2];
          if (use_kacc) {
            const double m = MdOrigInv[i];
            d0 = d0 - m * KAcc[x]; d1 = d1 - m * KAcc[x + 1]; d2 = d2 - m * KAcc[x + 2];
          }
          s0 += a0[j] * d0; s1 += a1[j] * d1; s2 += a2[j] * d2;
        }
      }
    }
  }
  s0 = group_sum<LPR>(s0); s1 = group_sum<LPR>(s1); s2 = group_sum<LPR>(s2);
  if (p < NNp && l == 0) R2[p] = (zeroP > 0 && p == zeroP) ? 0.0 : (s0 + s1) + s2;
}

int launch_step2_rhs(fsx_ctx* ctx, int iter) {
  // FP64 operands, 64^3: 16x8 0.307 ms, 16x4 0.310, 32x4 0.316, 16x2 0.312, 8x4 0.344, 8x8 0.346, 64x2 0.372,
  // 4x8 0.429
  constexpr int LPR = 16;
  const double o2 = 1.0000000000000000 / (ctx->dt * ctx->dt);
  const int use_kacc = iter != 1 ? 1 : 0;
  // step1's FP32 shadow of d, when it is of the current UnpHalf and KAcc term
  const bool drec = ctx->GTf && ctx->drec_ok && ctx->drec_kacc == (use_kacc == 1);
  const float4* dr = reinterpret_cast<const float4*>(ctx->Drec);
  // with the FP32 operand, 32 lanes x 1 load a row: 122.7 us at 64^3 (32x2
  // 131.6, 16x1 126.4, 16x2 134.3, 32x4 139.6, 16x4 142.8, 16x8 161.0)
  const bool cls = g_classes_on(ctx);
  const bool off = cls && ctx->GTc_base;   // the classes carry the column offsets
  // G^T as row classes with offsets, FP32 operand: 8 lanes x 4 entries a
  // round, 38.7 us at 64^3 (8x8 39.9, 16x4 40.2, 16x2 41.9, 32x4 50.2, 32x2
  // 53.4, 64x2 108.9)
  if (drec && off)
    hipLaunchKernelGGL((k_step2_rhs<8, 4, float, true, 2>), dim3(rows_grid(ctx->NNp, 8)), dim3(kBlock), 0, ctx->stream,
                       ctx->NNp, ctx->NN, ctx->GTrow, ctx->GTc_base, ctx->GTf, ctx->UnpHalf, ctx->MdOrigInv, ctx->KAcc,
                       use_kacc, o2, ctx->zeroP, ctx->R2, dr, ctx->GTc_info, ctx->GTc_val);
  // G^T as row classes with per-row columns, FP32 operand: 16 lanes x 4
  // entries a round, 54.2 us at 64^3 (16x2 55.8, 8x4 56.1, 32x2 64.1, 32x1
  // 76.1, 64x2 107.5)
  else if (drec && cls)
    hipLaunchKernelGGL((k_step2_rhs<16, 4, float, true, 1>), dim3(rows_grid(ctx->NNp, 16)), dim3(kBlock), 0,
                       ctx->stream, ctx->NNp, ctx->NN, ctx->GTrow, ctx->GTc_col, ctx->GTf, ctx->UnpHalf, ctx->MdOrigInv,
                       ctx->KAcc, use_kacc, o2, ctx->zeroP, ctx->R2, dr, ctx->GTc_info, ctx->GTc_val);
  else if (drec)
    hipLaunchKernelGGL((k_step2_rhs<32, 1, float, true>), dim3(rows_grid(ctx->NNp, 32)), dim3(kBlock), 0, ctx->stream,
                       ctx->NNp, ctx->NN, ctx->GTrow, ctx->GTcol, ctx->GTf, ctx->UnpHalf, ctx->MdOrigInv, ctx->KAcc,
                       use_kacc, o2, ctx->zeroP, ctx->R2, dr);
  else if (off)
    hipLaunchKernelGGL((k_step2_rhs<LPR, 8, float, false, 2>), dim3(rows_grid(ctx->NNp, LPR)), dim3(kBlock), 0,
                       ctx->stream, ctx->NNp, ctx->NN, ctx->GTrow, ctx->GTc_base, ctx->GTf, ctx->UnpHalf, ctx->MdOrigInv,
                       ctx->KAcc, use_kacc, o2, ctx->zeroP, ctx->R2, dr, ctx->GTc_info, ctx->GTc_val);
  else if (cls)
    hipLaunchKernelGGL((k_step2_rhs<LPR, 8, float, false, 1>), dim3(rows_grid(ctx->NNp, LPR)), dim3(kBlock), 0,
                       ctx->stream, ctx->NNp, ctx->NN, ctx->GTrow, ctx->GTc_col, ctx->GTf, ctx->UnpHalf, ctx->MdOrigInv,
                       ctx->KAcc, use_kacc, o2, ctx->zeroP, ctx->R2, dr, ctx->GTc_info, ctx->GTc_val);
  else if (ctx->GTf)
    hipLaunchKernelGGL((k_step2_rhs<LPR, 8, float, false>), dim3(rows_grid(ctx->NNp, LPR)), dim3(kBlock), 0, ctx->stream,
                       ctx->NNp, ctx->NN, ctx->GTrow, ctx->GTcol, ctx->GTf, ctx->UnpHalf, ctx->MdOrigInv, ctx->KAcc,
                       use_kacc, o2, ctx->zeroP, ctx->R2, dr);
  else
    hipLaunchKernelGGL((k_step2_rhs<LPR, 8, double, false>), dim3(rows_grid(ctx->NNp, LPR)), dim3(kBlock), 0, ctx->stream,
                       ctx->NNp, ctx->NN, ctx->GTrow, ctx->GTcol, ctx->GT, ctx->UnpHalf, ctx->MdOrigInv, ctx->KAcc,
                       use_kacc, o2, ctx->zeroP, ctx->R2, dr);
  return launch_check(ctx, "k_step2_rhs");
}

// ------------------------------------------------------------ step reductions
// Pnp1 = Pn + dt Pdot (BCH:3891-3893) and the partial pressure norms of the
// convergence check (BCH:2950-2957).
__global__ __launch_bounds__(kBlock) void k_pnp1(int NNp, const double* __restrict__ Pn, const double* __restrict__ Pdot,
                                                 const double* __restrict__ Pprev, double* __restrict__ Pnp1, double dt,
                                                 double* __restrict__ partials) {
  double v[2] = {0.0, 0.0};
#pragma unroll 4   // independent elements: their loads in flight together
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < NNp; i += kRedGrid * kBlock) {
    const double p = Pn[i] + dt * Pdot[i];
    Pnp1[i] = p;
    const double d = p - Pprev[i];
    v[0] += p * p; v[1] += d * d;
  }
  write_partials<2>(v, partials);
}

// velocity norms of the convergence check (BCH:2936-2943)
__global__ __launch_bounds__(kBlock) void k_vnorms(long long n, const double* __restrict__ u, const double* __restrict__ uprev,
                                                   double* __restrict__ partials) {
  double v[2] = {0.0, 0.0};
#pragma unroll 4   // independent elements: their loads in flight together
  for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n; i += (long long)kRedGrid * kBlock) {
    const double a = u[i], d = a - uprev[i];
    v[0] += a * a; v[1] += d * d;
  }
  write_partials<2>(v, partials);
}

// sums 2 partial arrays into dst[0], dst[1] (one block)
__global__ __launch_bounds__(kBlock) void k_finish2(const double* __restrict__ partials, double* __restrict__ dst) {
  double out[2];
  read_partials<2>(partials, out);
  if (threadIdx.x == 0) { dst[0] = out[0]; dst[1] = out[1]; }
}

int launch_pnp1(fsx_ctx* ctx) {
  hipLaunchKernelGGL(k_pnp1, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, ctx->NNp, ctx->Pn, ctx->Pdot, ctx->Pnp1_prev,
                     ctx->Pnp1, ctx->dt, ctx->partials);
  hipLaunchKernelGGL(k_finish2, dim3(1), dim3(kBlock), 0, ctx->stream, ctx->partials, &ctx->st->pp);
  return launch_check(ctx, "k_pnp1");
}

// ------------------------------------------------------------ step 3
// R3_c = -dt G_c Pdot - [iter>1] dt KAcc_c; BC rows 0; Acc = R3 MdInv;
// Unp1 = UnpHalf + dt Acc (BCH:3912-3974).
template <int LPR, int U, class VT, bool CLS = false>
__global__ __launch_bounds__(kBlock) void k_step3(int NN, const int* __restrict__ Grow, const int* __restrict__ Gcol,
                                                  const VT* __restrict__ G, const double* __restrict__ Pdot,
                                                  const double* __restrict__ KAcc, int use_kacc,
                                                  const unsigned char* __restrict__ bc, const double* __restrict__ MdInv,
                                                  const double* __restrict__ UnpHalf, double* __restrict__ Acc,
                                                  double* __restrict__ Unp1, double dt, double* __restrict__ Rout,
                                                  float* __restrict__ Aout, const double* __restrict__ GPn,
                                                  double* __restrict__ GPnext, const int2* __restrict__ Gc_row = nullptr,
                                                  const float4* __restrict__ Gc_ent = nullptr) {
  int l;
  const long long r = my_row<LPR>(&l);
  const bool out = r < NN && l < 3;
  const size_t i = 3 * (size_t)r + l;
  // the epilogue's operands, loaded before the row's gather so that their
  // latency overlaps it
  double gpn = 0.0, kacc = 0.0, uh = 0.0, m = 0.0;
  bool isbc = false;
  if (out) {
    if (GPnext) gpn = GPn[i];
    if (use_kacc) kacc = KAcc[i];
    isbc = bc[r] != 0;
    if (!Rout) { m = MdInv[r]; uh = UnpHalf[i]; }
  }
  double g0 = 0, g1 = 0, g2 = 0;
  if (CLS && r < NN) {
    // class form: entry k of the row is class entry start + k, at column
    // base + the entry's offset; the same lanes take the same entries as
    // the CSR loop below
    const int2 info = Gc_row[r];
    const int cs = info.y >> 8, len = info.y & 255;
    for (int base = l; base < len; base += U * LPR) {
      float4 e[U];
#pragma unroll
      for (int j = 0; j < U; j++) {
        const int k = base + j * LPR;
        e[j] = Gc_ent[cs + (k < len ? k : len - 1)];
        if (k >= len) e[j].y = e[j].z = e[j].w = 0.0f;
      }
#pragma unroll
      for (int j = 0; j < U; j++) {
        const double p = Pdot[info.x + __float_as_int(e[j].x)];
        g0 += (double)e[j].y * p; g1 += (double)e[j].z * p; g2 += (double)e[j].w * p;
      }
    }
  } else if (r < NN) {
    const int rs = Grow[r], re = Grow[r + 1];
    for (int base = rs + l; base < re; base += U * LPR) {
      int c[U];
      double a0[U], a1[U], a2[U];
#pragma unroll
      for (int j = 0; j < U; j++) {
        const int k = base + j * LPR, kk = k < re ? k : re - 1;
        c[j] = Gcol[kk];
        const bool ok = k < re;
        a0[j] = ok ? (double)G[3 * (size_t)kk] : 0.0;
        a1[j] = ok ? (double)G[3 * (size_t)kk + 1] : 0.0;
        a2[j] = ok ? (double)G[3 * (size_t)kk + 2] : 0.0;
      }
#pragma unroll
      for (int j = 0; j < U; j++) {
        const double p = Pdot[c[j]];
        g0 += a0[j] * p; g1 += a1[j] * p; g2 += a2[j] * p;
      }
    }
  }
  g0 = group_sum<LPR>(g0); g1 = group_sum<LPR>(g1); g2 = group_sum<LPR>(g2);
  if (out) {
    const double gg = l == 0 ? g0 : (l == 1 ? g1 : g2);
    // G Pnp1 = G Pn + dt G Pdot (Pnp1 = Pn + dt Pdot, k_pnp1): the next
    // time step's G Pn when this is the step's last inner iteration
    if (GPnext) GPnext[i] = gpn + dt * gg;
    double R = -dt * gg;
    if (use_kacc) R = R - dt * kacc;
    if (isbc) R = 0.0;
    if (Rout) { Rout[i] = R; return; }   // consistent mass: the caller solves for Acc
    const double a = R * m;
    Acc[i] = a;
    if (Aout) Aout[8 * (size_t)r + l] = (float)a;   // the FP32 shadow step1 reads
    Unp1[i] = uh + dt * a;
  }
}

static void launch_step3_rows(fsx_ctx* ctx, int iter, double* Rout = nullptr) {
  // FP64 G values, 64^3: 8x2 0.253 ms, 8x4 0.266, 4x2 0.290, 4x8 0.298, 4x4 0.307, 16x2 0.312, 8x8 0.319,
  // 16x4 0.343
  constexpr int LPR = 8;
  // Acc's FP32 shadow, kept where step1 reads it (lumped mass, FP32 matrices)
  float* aout = (ctx->Gf && ctx->rec_mem && !Rout) ? ctx->Arec : nullptr;
  // with the element form of step1 (FP32 G, lumped mass), G Pnp1 is carried
  // to the next time step from the G Pdot formed here instead of a pass over
  // G at its first step1
  const bool carry = ctx->Gf && !Rout && ctx->gpn_valid && ctx->use_step1_elements && ctx->GPn_next;
  const double* gpn = carry ? ctx->GPn : nullptr;
  double* gnext = carry ? ctx->GPn_next : nullptr;
  // FP32 G values: 4 lanes x 1 load a row, 225.7 us at 64^3 (4x2 237.5, 8x1
  // 245.9, 8x2 252.6, 4x3 254.6, 8x4 261.7, 16x1 347.0)
  // G as row classes: 4 lanes x 2 entries a round, 142.5 us at 64^3 (4x3
  // 142.0, 4x4 147.9, 4x1 153.8, 8x2 191.1, 8x1 199.7)
  if (g_classes_on(ctx))
    hipLaunchKernelGGL((k_step3<4, 2, float, true>), dim3(rows_grid(ctx->NN, 4)), dim3(kBlock), 0, ctx->stream, ctx->NN,
                       ctx->Grow, ctx->Gcol, ctx->Gf, ctx->Pdot, ctx->KAcc, iter != 1 ? 1 : 0, ctx->bcflag, ctx->MdInv,
                       ctx->UnpHalf, ctx->Acc, ctx->Unp1, ctx->dt, Rout, aout, gpn, gnext, ctx->Gc_row, ctx->Gc_ent);
  else if (ctx->Gf)
    hipLaunchKernelGGL((k_step3<4, 1, float>), dim3(rows_grid(ctx->NN, 4)), dim3(kBlock), 0, ctx->stream, ctx->NN,
                       ctx->Grow, ctx->Gcol, ctx->Gf, ctx->Pdot, ctx->KAcc, iter != 1 ? 1 : 0, ctx->bcflag, ctx->MdInv,
                       ctx->UnpHalf, ctx->Acc, ctx->Unp1, ctx->dt, Rout, aout, gpn, gnext);
  else
    hipLaunchKernelGGL((k_step3<LPR, 2, double>), dim3(rows_grid(ctx->NN, LPR)), dim3(kBlock), 0, ctx->stream, ctx->NN,
                       ctx->Grow, ctx->Gcol, ctx->G, ctx->Pdot, ctx->KAcc, iter != 1 ? 1 : 0, ctx->bcflag, ctx->MdInv,
                       ctx->UnpHalf, ctx->Acc, ctx->Unp1, ctx->dt, Rout, aout, nullptr, nullptr);
  ctx->arec_ok = aout != nullptr;
  ctx->gpn_next_valid = carry;
}

int launch_step3(fsx_ctx* ctx, int iter) {
  if (ctx->mass_mode != FSX_MASS_LUMPED) {
    launch_step3_rows(ctx, iter, ctx->mass_b);
    FSX_TRY(launch_check(ctx, "k_step3"));
    const int rc = mass_update_step3(ctx);
    if (rc != FSX_OK && rc != FSX_ENOCONV) return rc;
  } else {
    launch_step3_rows(ctx, iter);
  }
  hipLaunchKernelGGL(k_vnorms, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, 3LL * ctx->NN, ctx->Unp1, ctx->Unp1_prev,
                     ctx->partials);
  hipLaunchKernelGGL(k_finish2, dim3(1), dim3(kBlock), 0, ctx->stream, ctx->partials, &ctx->st->uu);
  return launch_check(ctx, "k_step3");
}

// max |Unp1 - Un| / dt over 3NN (BCH:3048-3057)
__global__ __launch_bounds__(kBlock) void k_maxacc(long long n, const double* __restrict__ Unp1, const double* __restrict__ Un,
                                                   double oneOverdt, double* __restrict__ partials) {
  double m = 0.0;
#pragma unroll 4   // independent elements: their loads in flight together
  for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n; i += (long long)kRedGrid * kBlock)
    m = fmax(m, fabs((Unp1[i] - Un[i]) * oneOverdt));
  m = block_max_all(m);
  if (threadIdx.x == 0) partials[blockIdx.x] = m;
}

__global__ __launch_bounds__(kBlock) void k_maxacc_finish(const double* __restrict__ partials, StepScalars* st) {
  const double m = block_max_all(partials[threadIdx.x]);
  if (threadIdx.x == 0) st->maxacc = m;
}

int launch_maxacc(fsx_ctx* ctx) {
  hipLaunchKernelGGL(k_maxacc, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, 3LL * ctx->NN, ctx->Unp1, ctx->Un,
                     1.0000000000000000 / ctx->dt, ctx->partials);
  hipLaunchKernelGGL(k_maxacc_finish, dim3(1), dim3(kBlock), 0, ctx->stream, ctx->partials, ctx->st);
  return launch_check(ctx, "k_maxacc");
}

// ------------------------------------------------------------ generic SpMVs (tooling)
// y = alpha * A x + beta * y for one CSR value array with value stride `vs`
// (G components are interleaved by 3).
template <int LPR>
__global__ __launch_bounds__(kBlock) void k_csrmv(int nrows, const int* __restrict__ rp, const int* __restrict__ ci,
                                                  const double* __restrict__ v, int vs, const double* __restrict__ x,
                                                  double* __restrict__ y, double alpha, double beta) {
  int l;
  const long long r = my_row<LPR>(&l);
  double s = 0;
  if (r < nrows)
    for (int k = rp[r] + l; k < rp[r + 1]; k += LPR) s += v[(size_t)vs * k] * x[ci[k]];
  s = group_sum<LPR>(s);
  if (r < nrows && l == 0) y[r] = (beta == 0.0 ? 0.0 : beta * y[r]) + alpha * s;
}

// ------------------------------------------------------------ Jacobi-PCG
// Textbook preconditioned CG with x0 = 0 (BCH:3999-4004), Jacobi
// preconditioner (ipar[10] = 1, BCH:4017) and the stop test
// ||r||_2 <= rtol ||b||_2 after each update. Iteration j is four kernels:
//   K1 q = Z p                      (the SpMV; the dominant kernel)
//   K2 partials of p.q
//   K3 alpha = rz_j / pq; x += alpha p; r -= alpha q; z = dinv r;
//      partials of r.r and r.z
//   K4 beta = rz_{j+1} / rz_j; p = z + beta p; block 0 records the stop test
// Scalars live on the device (SolveScalars); launches after a stop do
// nothing, so the host only polls between chunks of iterations. T is the type
// of Z and the Krylov vectors (double, or float for the mixed solver); all
// dot products accumulate in FP64.
template <class T> struct PcgCfg;
template <> struct PcgCfg<double> { static constexpr int LPR = 16, U = 8; };
template <> struct PcgCfg<float> { static constexpr int LPR = 32, U = 4; };

// pq[b] = the block's sum of v over its lanes, wavefronts in order (every
// wavefront of the block calls it once)
__device__ __forceinline__ void block_pq(double v, int b, double* __restrict__ pq) {
  __shared__ double wsum[kBlock / kWave];
  v = group_sum<kWave>(v);
  if ((threadIdx.x & (kWave - 1)) == 0) wsum[threadIdx.x / kWave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < kBlock / kWave; i++) t += wsum[i];
    pq[b] = t;
  }
}

template <class T, int LPR, int U, bool PAT>
__global__ __launch_bounds__(kBlock) void k_pcg_spmv(int n, const int* __restrict__ rp, ColSrc cs,
                                                     const T* __restrict__ val, const T* __restrict__ p, T* __restrict__ q,
                                                     const SolveScalars* __restrict__ sc, double* __restrict__ pq) {
  if (sc && sc->stop) return;
  int l;
  const long long r = my_row<LPR>(&l);
  T s = 0;
  if (r < n) {
    const int rs = rp[r], re = rp[r + 1];
    const int cb = col_base<PAT>(cs, r, rs);
    for (int base = rs + l; base < re; base += U * LPR) {
      int c[U];
      T a[U];
#pragma unroll
      for (int j = 0; j < U; j++) {
        const int k = base + j * LPR, kk = k < re ? k : re - 1;
        c[j] = col_at<PAT>(cs, r, cb, kk);
        a[j] = k < re ? val[kk] : T(0);
      }
#pragma unroll
      for (int j = 0; j < U; j++) s += a[j] * p[c[j]];
    }
  }
  if constexpr (sizeof(T) == 8) s = group_sum<LPR>(s); else s = group_sumf<LPR>(s);
  if (r < n && l == 0) q[r] = s;
  // with pq: the block's p.q partial (its rows in order), so that no
  // separate dot pass reads p and q again
  if (pq) block_pq(r < n && l == 0 ? (double)p[r] * (double)s : 0.0, blockIdx.x, pq);
}


// q = Z p on the FP32 row classes: row r sums the entries of its class,
// (offset o, value a) pairs, as a * p[r + o], in entry order. The class table
// is a few tens of KB and stays in cache, so a launch moves little more than
// the row list, p and q. One wavefront per chunk of up to 64 rows of one
// class: the entries are wave-uniform and come through scalar loads, and
// entry k's gathers p[r + o_k] follow the chunk's rows (contiguous for the
// interior of a structured grid). With XCD, consecutive blocks of chunks go
// to one XCD (blocks are dealt to the 8 XCDs round robin), so each XCD's L2
// holds one slab of p instead of all of it; on the 64^3 cavity the launch
// order measured faster (9.05 us against 9.4 with the remap; U = 4, 16, 32
// with the remap: 10.1, 9.3, 10.1), so the solvers launch it without.
// With pq non-null each block also writes its chunks' p.q (FP64 sum of FP32
// products, its wavefronts in order) to pq[block]: the solvers take p.Zp
// from these instead of a separate dot pass.
template <int U, bool XCD, int S>
__global__ __launch_bounds__(kBlock) void k_zcls_spmv(int nchunk, const int4* __restrict__ chunk,
                                                      const int* __restrict__ rows, const int2* __restrict__ ent,
                                                      const float* __restrict__ p, float* __restrict__ q,
                                                      const SolveScalars* __restrict__ sc, double* __restrict__ pq) {
  static_assert((kBlock / kWave) % S == 0, "a chunk's wavefronts share a block");
  if (sc && sc->stop) return;
  int b = blockIdx.x;
  if constexpr (XCD) b = (b % 8) * (gridDim.x / 8) + b / 8;   // gridDim.x is a multiple of 8
  const int wv = threadIdx.x / kWave;
  const int wg = __builtin_amdgcn_readfirstlane(b * (kBlock / kWave) + wv);
  const int w = wg / S, h = wg % S;   // chunk, and this wavefront's share of its entries
  const int lane = threadIdx.x & (kWave - 1);
  float s = 0.0f;
  int r = 0, len = 0;
  if (w < nchunk) {
    const int4 c = chunk[w];
    const int rs = __builtin_amdgcn_readfirstlane(c.x);
    len = __builtin_amdgcn_readfirstlane(c.y);
    const int es = __builtin_amdgcn_readfirstlane(c.z), el = __builtin_amdgcn_readfirstlane(c.w);
    r = rows[rs + (lane < len ? lane : 0)];
    const int2* __restrict__ e = ent + es;
    const int k1 = el * (h + 1) / S;
    int k = el * h / S;
    for (; k + U <= k1; k += U) {
      int2 t[U];
      float v[U];
#pragma unroll
      for (int j = 0; j < U; j++) t[j] = e[k + j];
#pragma unroll
      for (int j = 0; j < U; j++) v[j] = p[r + t[j].x];
#pragma unroll
      for (int j = 0; j < U; j++) s += __int_as_float(t[j].y) * v[j];
    }
    for (; k < k1; k++) {
      const int2 t = e[k];
      s += __int_as_float(t.y) * p[r + t.x];
    }
  }
  if constexpr (S > 1) {
    // the chunk's S partial sums, added in wavefront order by its first one
    __shared__ float part[kBlock / kWave][kWave];
    part[wv][lane] = s;
    __syncthreads();
    if (h == 0)
      for (int j = 1; j < S; j++) s += part[wv + j][lane];
  }
  const bool own = h == 0 && w < nchunk && lane < len;
  if (own) q[r] = s;
  if (pq) block_pq(own ? (double)p[r] * (double)s : 0.0, b, pq);
}

// The solvers' Z (structural zeros dropped): FP64 rows through the pattern
// dictionary when there is one
static ColSrc zr_cols(fsx_ctx* ctx) { return ColSrc{ctx->Zr_col, ctx->Zr_pat, ctx->Zr_off}; }

constexpr int kClsU = 8;
constexpr int kClsS = 2;   // wavefronts per chunk of the row-class SpMV

// blocks of the row-class SpMV (a multiple of 8, one p.q partial each)
static unsigned zcls_blocks(const fsx_ctx* ctx) {
  const long long wpb = kBlock / kWave;
  return (unsigned)(((long long)ctx->Zc_nchunk * kClsS + 8 * wpb - 1) / (8 * wpb) * 8);
}

// q = Z p: FP64 on the solvers' rows; FP32 on the row classes, or on the
// FP32 rounding of the same rows when they form no classes
template <class T>
static void launch_zspmv(fsx_ctx* ctx, const T* p, T* q, const SolveScalars* sc, double* pq = nullptr,
                         int* npq = nullptr) {
  const int n = ctx->NNp;
  if constexpr (sizeof(T) == 4) {
    if (ctx->Zc_chunk) {
      hipLaunchKernelGGL((k_zcls_spmv<kClsU, false, kClsS>), dim3(zcls_blocks(ctx)), dim3(kBlock), 0, ctx->stream,
                         ctx->Zc_nchunk,
                         ctx->Zc_chunk, ctx->Zc_rows, ctx->Zc_ent, p, q, sc, nullptr);
      return;
    }
  }
  constexpr int LPR = PcgCfg<T>::LPR, U = PcgCfg<T>::U;
  // both precisions on the rows of the FP64 form (structural zeros dropped)
  const bool fp64 = sizeof(T) == 8;
  const ColSrc cs = zr_cols(ctx);
  const int* rp = ctx->Zr_row;
  const T* val = fp64 ? (const T*)(const void*)ctx->Zr_val : (const T*)(const void*)ctx->Zr_valf;
  // the dictionary form uses 8 lanes x 16 loads per row in both precisions
  // (a sweep of 8x16, 16x8, 32x4, 64x2 on the 64^3 Z: 36.7, 56.2, 64.3,
  // 82.0 us in FP32; 44.5, 45.5, 53.4, 80.9 us in FP64): more independent
  // loads in flight per lane hide the per-row pattern lookup
  if (cs.row_pat)
    hipLaunchKernelGGL((k_pcg_spmv<T, 8, 16, true>), dim3(rows_grid(n, 8)), dim3(kBlock), 0, ctx->stream, n, rp, cs, val,
                       p, q, sc, pq);
  else
    hipLaunchKernelGGL((k_pcg_spmv<T, LPR, U, false>), dim3(rows_grid(n, LPR)), dim3(kBlock), 0, ctx->stream, n, rp, cs,
                       val, p, q, sc, pq);
  if (npq) *npq = (int)(cs.row_pat ? rows_grid(n, 8) : rows_grid(n, LPR));
}

__global__ void k_to_f32(int n, const double* __restrict__ a, float* __restrict__ b) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) b[i] = (float)a[i];
}
__global__ void k_to_f64(int n, const float* __restrict__ a, double* __restrict__ b) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) b[i] = (double)a[i];
}

int spmv_host_op(fsx_ctx* ctx, int which, const double* dx, double* dy, double alpha, double beta) {
  constexpr int LPR = 16;
  if (which == 0) {
    hipLaunchKernelGGL(k_csrmv<LPR>, dim3(rows_grid(ctx->NN, LPR)), dim3(kBlock), 0, ctx->stream, ctx->NN, ctx->Mrow,
                       ctx->Mcol, ctx->K, 1, dx, dy, alpha, beta);
  } else if (which >= 1 && which <= 3) {
    hipLaunchKernelGGL(k_csrmv<LPR>, dim3(rows_grid(ctx->NN, LPR)), dim3(kBlock), 0, ctx->stream, ctx->NN, ctx->Grow,
                       ctx->Gcol, ctx->G + (which - 1), 3, dx, dy, alpha, beta);
  } else if (which == 4) {
    hipLaunchKernelGGL(k_csrmv<LPR>, dim3(rows_grid(ctx->NNp, LPR)), dim3(kBlock), 0, ctx->stream, ctx->NNp, ctx->Zrow,
                       ctx->Zcol, ctx->Z, 1, dx, dy, alpha, beta);
  } else if (which == 8) {
    launch_zspmv<double>(ctx, dx, dy, nullptr);
  } else if (which == 9) {
    hipLaunchKernelGGL(k_to_f32, dim3(ew_grid(ctx->NNp)), dim3(kBlock), 0, ctx->stream, ctx->NNp, dx, ctx->fp);
    launch_zspmv<float>(ctx, ctx->fp, ctx->fq, nullptr);
    hipLaunchKernelGGL(k_to_f64, dim3(ew_grid(ctx->NNp)), dim3(kBlock), 0, ctx->stream, ctx->NNp, ctx->fq, dy);
  } else {
    hipLaunchKernelGGL(k_csrmv<LPR>, dim3(rows_grid(ctx->NNp, LPR)), dim3(kBlock), 0, ctx->stream, ctx->NNp, ctx->GTrow,
                       ctx->GTcol, ctx->GT + (which - 5), 3, dx, dy, alpha, beta);
  }
  return launch_check(ctx, "k_csrmv");
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_pcg_init(int n, const double* __restrict__ b, const T* __restrict__ dinv,
                                                     T* __restrict__ x, T* __restrict__ r, T* __restrict__ z,
                                                     T* __restrict__ p, double* __restrict__ partials) {
  double v[2] = {0.0, 0.0};
#pragma unroll 4   // independent elements: their loads in flight together
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += kRedGrid * kBlock) {
    const T bi = (T)b[i];
    const T zi = dinv[i] * bi;
    x[i] = (T)0; r[i] = bi; z[i] = zi; p[i] = zi;
    v[0] += (double)bi * (double)zi;
    v[1] += (double)bi * (double)bi;
  }
  write_partials<2>(v, partials);
}

__global__ __launch_bounds__(kBlock) void k_pcg_init_finish(const double* __restrict__ partials, SolveScalars* sc,
                                                            double rtol, int maxit) {
  double out[2];
  read_partials<2>(partials, out);
  if (threadIdx.x == 0) {
    sc->rz_v[0] = out[0]; sc->rz_v[1] = 0.0;
    sc->bb = out[1]; sc->rr = out[1]; sc->rr_anchor = out[1]; sc->rr_hi = out[1];
    sc->rtol2bb = rtol * rtol * out[1];
    sc->iters = 0; sc->maxit = maxit;
    sc->stop = (out[1] == 0.0) ? 1 : 0;
    sc->active_v[0] = 0; sc->active_v[1] = 0;
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_pcg_dot(int n, const T* __restrict__ p, const T* __restrict__ q,
                                                    double* __restrict__ partials, const SolveScalars* __restrict__ sc) {
  if (sc->stop) return;
  double v[1] = {0.0};
#pragma unroll 4   // independent elements: their loads in flight together
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += kRedGrid * kBlock) v[0] += (double)p[i] * (double)q[i];
  write_partials<1>(v, partials);
}

template <class T>
// p.q from npq partial sums: k_pcg_dot's kRedGrid (slot 0 of partials) or
// the FP32 row-class SpMV's one per chunk
__global__ __launch_bounds__(kBlock) void k_pcg_update(int j, int n, T* __restrict__ x, T* __restrict__ r,
                                                       T* __restrict__ z, const T* __restrict__ p, const T* __restrict__ q,
                                                       const T* __restrict__ dinv, double* __restrict__ partials,
                                                       SolveScalars* sc, const double* __restrict__ pqp, int npq) {
  const int stop = sc->stop;
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->active_v[j & 1] = stop ? 0 : 1;
  if (stop) return;
  const double pq = sum_parts(pqp, npq);
  const T a = (T)(sc->rz_v[j & 1] / pq);
  double v[2] = {0.0, 0.0};
#pragma unroll 4   // independent elements: their loads in flight together
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += kRedGrid * kBlock) {
    x[i] += a * p[i];
    const T ri = r[i] - a * q[i];
    r[i] = ri;
    const T zi = dinv[i] * ri;
    z[i] = zi;
    v[0] += (double)ri * (double)ri;
    v[1] += (double)ri * (double)zi;
  }
  write_partials<2>(v, partials + kRedGrid);
}

// KIND 1 (FP32 inner solve of the mixed solver): the recurrence residual
// reaching the target (stop 1) or dropping by delta^2 since the last FP64
// residual (stop 5) pauses for a reliable update by the host.
template <class T, int KIND>
__global__ __launch_bounds__(kBlock) void k_pcg_direction(int j, int n, const T* __restrict__ z, T* __restrict__ p,
                                                          const double* __restrict__ partials, double delta2,
                                                          SolveScalars* sc) {
  if (!sc->active_v[j & 1]) return;
  double s[2];
  read_partials<2>(partials + kRedGrid, s);
  const double rz_old = sc->rz_v[j & 1];
  const bool bad = !(s[0] == s[0]) || !(s[1] == s[1]) || !(rz_old == rz_old) || rz_old == 0.0;
  const T be = bad ? T(0) : (T)(s[1] / rz_old);
#pragma unroll 4   // independent elements: their loads in flight together
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += kRedGrid * kBlock) p[i] = z[i] + be * p[i];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const int it = sc->iters + 1;
    sc->iters = it;
    sc->rr = s[0];
    sc->rz_v[(j + 1) & 1] = s[1];
    if (bad) sc->stop = 3;
    else if (s[0] <= sc->rtol2bb) sc->stop = 1;
    else if (KIND == 1 && s[0] <= delta2 * sc->rr_anchor) sc->stop = 5;
    else if (it >= sc->maxit) sc->stop = 2;
  }
}

int launch_pcg_spmv_pq_f32(fsx_ctx* ctx, const float* p, float* q, const SolveScalars* sc, double* pq, int* npq);

// The FP32 SpMV's p.q partial sums (one per row-class chunk, or per block of
// a 256-block grid); the captured solver iterations hold the buffer, so they
// are dropped when it is reallocated.
int ensure_fpq(fsx_ctx* ctx) {
  // >= the SpMV's block count: row-class chunks, or CSR rows (32 lanes a row at most)
  const int need = std::max({ctx->Zc_chunk ? (int)zcls_blocks(ctx) : 0, kRedGrid, (int)rows_grid(ctx->NNp, 32)});
  if (ctx->fpq_n >= need) return FSX_OK;
  if (ctx->fpq) dev_free(ctx, ctx->fpq);
  ctx->fpq = nullptr;
  ctx->fpq_n = 0;
  FSX_TRY(alloc(ctx, &ctx->fpq, (size_t)need));
  ctx->fpq_n = need;
  for (IterGraph* g : {&ctx->graph_fp64, &ctx->graph_mixed, &ctx->graph_learned}) iter_graph_reset(*g);
  return FSX_OK;
}

// One iteration on the solve scalars sc and the reduction partials
// (3 kRedGrid doubles); the solvers pass the context's own, the
// allocation-free solve its scratch.
template <class T>
static void pcg_launch_iteration(fsx_ctx* ctx, int j, const T* dinv, T* x, T* r, T* z, T* p, T* q,
                                 int kind, double delta2, SolveScalars* sc, double* partials) {
  constexpr int LPR = PcgCfg<T>::LPR, U = PcgCfg<T>::U;
  const int n = ctx->NNp;
  const double* pqp = partials;
  int npq = kRedGrid;
  bool fused = false;
  if constexpr (sizeof(T) == 4) {
    if (ctx->fpq) {   // the FP32 SpMV (row classes or CSR rows) writes p.q per block: no separate dot pass
      (void)launch_pcg_spmv_pq_f32(ctx, p, q, sc, ctx->fpq, &npq);
      pqp = ctx->fpq;
      fused = true;
    }
  }
  if (!fused) {
    launch_zspmv<T>(ctx, p, q, sc);
    hipLaunchKernelGGL(k_pcg_dot<T>, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, n, p, q, partials, sc);
  }
  hipLaunchKernelGGL(k_pcg_update<T>, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, j, n, x, r, z, p, q, dinv,
                     partials, sc, pqp, npq);
  if (kind == 0)
    hipLaunchKernelGGL((k_pcg_direction<T, 0>), dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, j, n, z, p, partials,
                       delta2, sc);
  else
    hipLaunchKernelGGL((k_pcg_direction<T, 1>), dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, j, n, z, p, partials,
                       delta2, sc);
}

// Enqueues iterations j0, j0+1, ... in chunks (the first sized from the
// previous solve) and polls the device scalars after each chunk until the
// stop flag is set. Returns with *ctx->h_sc current.
template <class T>
static int pcg_iterate(fsx_ctx* ctx, int j0, const T* dinv, T* x, T* r, T* z, T* p, T* q, int kind,
                       double delta2, int first_chunk, int maxit) {
  int j = j0;
  int chunk = std::max(4, first_chunk);
  IterGraph& G = sizeof(T) == 8 ? ctx->graph_fp64 : ctx->graph_mixed;
  const unsigned long long key = (unsigned long long)(uintptr_t)x;   // the only caller-owned buffer
  auto issue = [&](int jj) -> int {
    pcg_launch_iteration<T>(ctx, jj, dinv, x, r, z, p, q, kind, delta2, ctx->sc, ctx->partials);
    return FSX_OK;
  };
  for (;;) {
    int done = 0;
    FSX_TRY(run_iterations(ctx, G, key, j, chunk, &done, issue));
    j += done;
    FSX_TRY(read_scalars(ctx));
    if (ctx->h_sc->stop) return FSX_OK;
    if (j - j0 > maxit + 64) return fail(ctx, FSX_ESTATE, "PCG did not terminate");
    chunk = 4;
  }
}

static void timing_begin(fsx_ctx* ctx) {
  if (ctx->timing) (void)hipEventRecord(ctx->ev0, ctx->stream);
}
static void timing_end(fsx_ctx* ctx, int iters) {
  if (!ctx->timing) return;
  (void)hipEventRecord(ctx->ev1, ctx->stream);
  (void)hipEventSynchronize(ctx->ev1);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
  ctx->tm.pcg_ms += ms;
  ctx->tm.pcg_spmv_launches += iters;
  ctx->tm.pcg_iters += iters;
}

static int solve_fp64(fsx_ctx* ctx, const double* b, double* x, int* iters, double* relres) {
  const int n = ctx->NNp;
  hipLaunchKernelGGL(k_pcg_init<double>, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, n, b, ctx->Zdinv, x, ctx->wr,
                     ctx->wz, ctx->wp, ctx->partials);
  hipLaunchKernelGGL(k_pcg_init_finish, dim3(1), dim3(kBlock), 0, ctx->stream, ctx->partials, ctx->sc, ctx->rtol, ctx->maxit);
  FSX_TRY(launch_check(ctx, "k_pcg_init"));
  FSX_TRY(pcg_iterate<double>(ctx, 0, ctx->Zdinv, x, ctx->wr, ctx->wz, ctx->wp, ctx->wq, 0, 0.0,
                              segment_guess(std::min(ctx->last_iters, ctx->maxit), 1.0, ctx->rtol, 4), ctx->maxit));
  const SolveScalars& s = *ctx->h_sc;
  *iters = s.iters;
  *relres = s.bb > 0 ? std::sqrt(s.rr / s.bb) : 0.0;
  ctx->last_iters = s.iters;
  if (s.stop == 3) return fail(ctx, FSX_ENUMERIC, "PCG met NaN/Inf");
  if (s.stop == 2) { ctx->err = "PCG reached maxit"; return FSX_ENOCONV; }
  return FSX_OK;
}

// ------------------------------------------------------------ mixed precision
// FP32 Jacobi-PCG with reliable updates: the CG recurrence (Z, x, r, z, p, q)
// runs in FP32; whenever its residual has dropped by delta since the last
// anchor, or reaches the target, the FP32 iterate is folded into the FP64
// solution x_hi, the residual is recomputed in FP64 as b - Z x_hi (the
// refinement residual), the FP64 stop test ||b - Z x_hi|| <= rtol ||b|| is
// applied, and the FP32 residual is replaced by the rounded FP64 one. The
// search direction is kept, so the Krylov information is not lost as in a
// restarted refinement, and the iteration count stays that of FP64 CG.
__global__ __launch_bounds__(kBlock) void k_mx_fold(int n, double* __restrict__ xh, float* __restrict__ xl) {
#pragma unroll 4   // independent elements: their loads in flight together
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += kRedGrid * kBlock) {
    xh[i] += (double)xl[i];
    xl[i] = 0.0f;
  }
}

// rh = b - Z xh (FP64)
template <int LPR, int U, bool PAT>
__global__ __launch_bounds__(kBlock) void k_residual(int n, const int* __restrict__ rp, ColSrc cs,
                                                     const double* __restrict__ val, const double* __restrict__ xh,
                                                     const double* __restrict__ b, double* __restrict__ rh) {
  int l;
  const long long i = my_row<LPR>(&l);
  double s = 0;
  if (i < n) {
    const int rs = rp[i], re = rp[i + 1];
    const int cb = col_base<PAT>(cs, i, rs);
    for (int base = rs + l; base < re; base += U * LPR) {
      int c[U];
      double a[U];
#pragma unroll
      for (int j = 0; j < U; j++) {
        const int k = base + j * LPR, kk = k < re ? k : re - 1;
        c[j] = col_at<PAT>(cs, i, cb, kk);
        a[j] = k < re ? val[kk] : 0.0;
      }
#pragma unroll
      for (int j = 0; j < U; j++) s += a[j] * xh[c[j]];
    }
  }
  s = group_sum<LPR>(s);
  if (i < n && l == 0) rh[i] = b[i] - s;
}

// rh = b - Z xh in FP64 on the row classes: each value is the class's FP32
// value plus the row's FP32 correction (Z64 to within 2^-48 of each value),
// a wavefront per chunk of rows of one class, a lane per row
template <int U>
__global__ __launch_bounds__(kBlock) void k_zcls_residual(int nchunk, const int4* __restrict__ chunk,
                                                          const int* __restrict__ rows, const int2* __restrict__ ent,
                                                          const float* __restrict__ delta,
                                                          const long long* __restrict__ doff,
                                                          const double* __restrict__ xh, const double* __restrict__ b,
                                                          double* __restrict__ rh) {
  const int w = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  if (w >= nchunk) return;
  const int lane = threadIdx.x & (kWave - 1);
  const int4 c = chunk[w];
  const int rs = __builtin_amdgcn_readfirstlane(c.x), len = __builtin_amdgcn_readfirstlane(c.y);
  const int es = __builtin_amdgcn_readfirstlane(c.z), el = __builtin_amdgcn_readfirstlane(c.w);
  const int r = rows[rs + (lane < len ? lane : 0)];
  const int2* __restrict__ e = ent + es;
  const float* __restrict__ dl = delta + doff[w] + lane;
  double s = 0.0;
  int k = 0;
  for (; k + U <= el; k += U) {
    int2 t[U];
    double v[U];
    float d[U];
#pragma unroll
    for (int j = 0; j < U; j++) t[j] = e[k + j];
#pragma unroll
    for (int j = 0; j < U; j++) { v[j] = xh[r + t[j].x]; d[j] = dl[(k + j) * 64]; }
#pragma unroll
    for (int j = 0; j < U; j++) s += ((double)__int_as_float(t[j].y) + (double)d[j]) * v[j];
  }
  for (; k < el; k++) {
    const int2 t = e[k];
    s += ((double)__int_as_float(t.y) + (double)dl[k * 64]) * xh[r + t.x];
  }
  if (lane < len) rh[r] = b[r] - s;
}

// b - Z x in FP64 on a tensor-product box grid from Z's Kronecker form
// (Kx My Mz + Mx Ky Mz + Mx My Kz, FP64 axis tables; the pinned node's
// penalty on its diagonal): one row a thread, per (dy, dz) neighbour line
// the K and M sums along x, then their weights. The tables are zero for
// couplings past the grid's ends, so the loads clamp to the grid and carry
// no branches.
__global__ __launch_bounds__(kBlock) void k_zten_residual(int nx, int ny, int nz, const double* __restrict__ coef,
                                                          int zp, double pen, const double* __restrict__ x,
                                                          const double* __restrict__ b, double* __restrict__ r) {
  const int id = blockIdx.x * kBlock + threadIdx.x;
  const int nxy = nx * ny;
  if (id >= nxy * nz) return;
  const int k = id / nxy, rem = id - k * nxy, j = rem / nx, i = rem - j * nx;
  const double* cx = coef + 10 * i;
  const double* cy = coef + 10 * (nx + j);
  const double* cz = coef + 10 * (nx + ny + k);
  int xo[5];
#pragma unroll
  for (int d = 0; d < 5; d++) xo[d] = min(max(i + d - 2, 0), nx - 1);
  double acc = 0.0;
#pragma unroll
  for (int dz = 0; dz < 5; dz++) {
    const int kk = min(max(k + dz - 2, 0), nz - 1);
#pragma unroll
    for (int dy = 0; dy < 5; dy++) {
      const int jj = min(max(j + dy - 2, 0), ny - 1);
      const double* line = x + (size_t)kk * nxy + (size_t)jj * nx;
      double tk = 0.0, tm = 0.0;
#pragma unroll
      for (int dx = 0; dx < 5; dx++) {
        const double v = line[xo[dx]];
        tk += cx[dx] * v;
        tm += cx[5 + dx] * v;
      }
      acc += tk * (cy[5 + dy] * cz[5 + dz]) + tm * (cy[dy] * cz[5 + dz] + cy[5 + dy] * cz[dz]);
    }
  }
  if (id == zp) acc += pen * x[id];
  r[id] = b[id] - acc;
}

static void launch_zresidual(fsx_ctx* ctx, const double* xh, const double* b, double* rh) {
  constexpr int LPR = PcgCfg<double>::LPR, U = PcgCfg<double>::U;
  const ColSrc cs = zr_cols(ctx);
  const int n = ctx->NNp;
  // the mixed and learned solvers' residuals from Z's Kronecker form where
  // it holds (vectors only: 7 MB instead of the row classes' 100 MB)
  if (ctx->zt64 && (ctx->solver_mode == FSX_SOLVER_MIXED_IR || ctx->solver_mode == FSX_SOLVER_LEARNED)) {
    hipLaunchKernelGGL(k_zten_residual, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream, ctx->zt_n[0],
                       ctx->zt_n[1], ctx->zt_n[2], ctx->zt64, ctx->zt_zp, ctx->zt_pen, xh, b, rh);
    return;
  }
  // the mixed and learned solvers' residuals on the row classes with FP64
  // corrections (half the bytes of the FP64 values); the FP64 and direct
  // solvers keep the FP64 values
  if (ctx->Zc_delta && (ctx->solver_mode == FSX_SOLVER_MIXED_IR || ctx->solver_mode == FSX_SOLVER_LEARNED)) {
    const int wpb = kBlock / kWave;
    hipLaunchKernelGGL(k_zcls_residual<kClsU>, dim3((ctx->Zc_nchunk + wpb - 1) / wpb), dim3(kBlock), 0, ctx->stream,
                       ctx->Zc_nchunk, ctx->Zc_chunk, ctx->Zc_rows, ctx->Zc_ent, ctx->Zc_delta, ctx->Zc_doff, xh, b, rh);
    return;
  }
  if (cs.row_pat)
    hipLaunchKernelGGL((k_residual<8, 16, true>), dim3(rows_grid(n, 8)), dim3(kBlock), 0, ctx->stream, n, ctx->Zr_row, cs,
                       ctx->Zr_val, xh, b, rh);
  else
    hipLaunchKernelGGL((k_residual<LPR, U, false>), dim3(rows_grid(n, LPR)), dim3(kBlock), 0, ctx->stream, n, ctx->Zr_row,
                       cs, ctx->Zr_val, xh, b, rh);
}

__global__ __launch_bounds__(kBlock) void k_sumsq(int n, const double* __restrict__ a, double* __restrict__ partials) {
  double v[1] = {0.0};
#pragma unroll 4   // independent elements: their loads in flight together
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += kRedGrid * kBlock) v[0] += a[i] * a[i];
  write_partials<1>(v, partials);
}

// FP64 stop test on ||b - Z x_hi||^2, or a new anchor and resume
__global__ __launch_bounds__(kBlock) void k_mx_check(const double* __restrict__ partials, SolveScalars* sc) {
  double out[1];
  read_partials<1>(partials, out);
  if (threadIdx.x == 0) {
    sc->rr_hi = out[0];
    if (!(out[0] == out[0])) sc->stop = 3;
    else if (out[0] <= sc->rtol2bb) sc->stop = 1;
    else if (sc->iters >= sc->maxit) sc->stop = 2;
    else { sc->stop = 0; sc->rr_anchor = out[0]; }
  }
}

// r = float(rh), z = dinv r, partials of r.z
__global__ __launch_bounds__(kBlock) void k_mx_replace(int n, const double* __restrict__ rh, const float* __restrict__ dinv,
                                                       float* __restrict__ r, float* __restrict__ z,
                                                       double* __restrict__ partials) {
  double v[1] = {0.0};
#pragma unroll 4   // independent elements: their loads in flight together
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += kRedGrid * kBlock) {
    const float ri = (float)rh[i];
    const float zi = dinv[i] * ri;
    r[i] = ri; z[i] = zi;
    v[0] += (double)ri * (double)zi;
  }
  write_partials<1>(v, partials);
}

__global__ __launch_bounds__(kBlock) void k_mx_replace_finish(int j, const double* __restrict__ partials, SolveScalars* sc) {
  double out[1];
  read_partials<1>(partials, out);
  if (threadIdx.x == 0) sc->rz_v[j & 1] = out[0];
}

static int solve_mixed(fsx_ctx* ctx, const double* b, double* x, int* iters, double* relres) {
  const int n = ctx->NNp;
  // a reliable update every 1e-4 of FP32 residual reduction (1e-2 and 1e-3
  // gave the same iteration counts at 64^3 and on the bent duct, with more
  // FP64 residuals: mixed solve 117.3 / 115.8 against 112.9 ms over 20 steps)
  const double delta = 1e-4;
  FSX_TRY(ensure_fpq(ctx));
  FSX_TRY(launch_fill(ctx, x, 0.0, n));
  hipLaunchKernelGGL(k_pcg_init<float>, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, n, b, ctx->Zdinvf, ctx->fx, ctx->fr,
                     ctx->fz, ctx->fp, ctx->partials);
  hipLaunchKernelGGL(k_pcg_init_finish, dim3(1), dim3(kBlock), 0, ctx->stream, ctx->partials, ctx->sc, ctx->rtol, ctx->maxit);
  FSX_TRY(launch_check(ctx, "k_pcg_init<float>"));
  const int seg = segment_guess(ctx->last_iters, delta, ctx->rtol, 4);
  int j = 0, guess = seg, status = 0, updates = 0;
  for (;;) {
    FSX_TRY(pcg_iterate<float>(ctx, j, ctx->Zdinvf, ctx->fx, ctx->fr, ctx->fz, ctx->fp, ctx->fq, 1, delta * delta,
                               guess, ctx->maxit));
    const int stop = ctx->h_sc->stop;
    if (stop == 3) return fail(ctx, FSX_ENUMERIC, "mixed PCG met NaN/Inf");
    if (stop == 1 && ctx->h_sc->bb == 0.0) break;   // b = 0: x = 0
    j = ctx->h_sc->iters;                            // next iteration index
    // reliable update: fold the FP32 iterate into x, FP64 residual and stop test
    hipLaunchKernelGGL(k_mx_fold, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, n, x, ctx->fx);
    launch_zresidual(ctx, x, b, ctx->wh);
    hipLaunchKernelGGL(k_sumsq, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, n, ctx->wh, ctx->partials);
    hipLaunchKernelGGL(k_mx_check, dim3(1), dim3(kBlock), 0, ctx->stream, ctx->partials, ctx->sc);
    FSX_TRY(launch_check(ctx, "mixed reliable update"));
    FSX_TRY(read_scalars(ctx));
    updates++;
    if (ctx->h_sc->stop == 3) return fail(ctx, FSX_ENUMERIC, "mixed residual met NaN/Inf");
    if (ctx->h_sc->stop == 1) break;
    if (ctx->h_sc->stop == 2) { status = 2; break; }
    hipLaunchKernelGGL(k_mx_replace, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, n, ctx->wh, ctx->Zdinvf, ctx->fr, ctx->fz,
                       ctx->partials);
    hipLaunchKernelGGL(k_mx_replace_finish, dim3(1), dim3(kBlock), 0, ctx->stream, j, ctx->partials, ctx->sc);
    FSX_TRY(launch_check(ctx, "k_mx_replace"));
    guess = seg;
  }
  const SolveScalars& s = *ctx->h_sc;
  *iters = s.iters;
  *relres = s.bb > 0 ? std::sqrt(s.rr_hi / s.bb) : 0.0;
  ctx->last_iters = s.iters;
  ctx->last_updates = updates;
  if (status == 2) { ctx->err = "mixed-precision PCG reached maxit"; return FSX_ENOCONV; }
  return FSX_OK;
}

// ------------------------------------------------------------ launchers used by the learned solver
int solve_mixed_entry(fsx_ctx* ctx, const double* b, double* x, int* iters, double* relres) {
  return solve_mixed(ctx, b, x, iters, relres);
}
int solve_fp64_entry(fsx_ctx* ctx, const double* b, double* x, int* iters, double* relres) {
  return solve_fp64(ctx, b, x, iters, relres);
}

int launch_pcg_spmv_f32(fsx_ctx* ctx, const float* p, float* q, const SolveScalars* sc) {
  launch_zspmv<float>(ctx, p, q, sc);
  return FSX_OK;
}


// q = Z p in FP32 with the p.q partials in pq (*npq of them): from the
// row-class SpMV itself (one per chunk), or from a separate pass when Z has
// no row classes
int launch_pcg_spmv_pq_f32(fsx_ctx* ctx, const float* p, float* q, const SolveScalars* sc, double* pq, int* npq) {
  if (ctx->Zc_chunk) {
    const unsigned blocks = zcls_blocks(ctx);
    hipLaunchKernelGGL((k_zcls_spmv<kClsU, false, kClsS>), dim3(blocks), dim3(kBlock), 0, ctx->stream, ctx->Zc_nchunk,
                       ctx->Zc_chunk, ctx->Zc_rows, ctx->Zc_ent, p, q, sc, pq);
    *npq = (int)blocks;
    return FSX_OK;
  }
  launch_zspmv<float>(ctx, p, q, sc, pq, npq);   // the block partials of p.q from the CSR SpMV itself
  return FSX_OK;
}

int launch_pcg_dot_f32(fsx_ctx* ctx, const float* p, const float* q) {
  hipLaunchKernelGGL(k_pcg_dot<float>, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, ctx->NNp, p, q, ctx->partials, ctx->sc);
  return FSX_OK;
}

__global__ void k_add_d(int n, double* __restrict__ x, const double* __restrict__ d) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] += d[i];
}

int launch_add(fsx_ctx* ctx, double* x, const double* d, int n) {
  hipLaunchKernelGGL(k_add_d, dim3(ew_grid(n)), dim3(kBlock), 0, ctx->stream, n, x, d);
  return launch_check(ctx, "k_add_d");
}

// wh = b - Z x (FP64), ||wh||^2 and the FP64 stop test of k_mx_check; reads
// the scalars back. With fold, first x += x_lo (FP32 iterate) and x_lo = 0.
static int residual_check(fsx_ctx* ctx, double* x, const double* b, bool fold, int* stop, double* rr_hi) {
  const int n = ctx->NNp;
  if (fold) hipLaunchKernelGGL(k_mx_fold, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, n, x, ctx->fx);
  launch_zresidual(ctx, x, b, ctx->wh);
  hipLaunchKernelGGL(k_sumsq, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, n, ctx->wh, ctx->partials);
  hipLaunchKernelGGL(k_mx_check, dim3(1), dim3(kBlock), 0, ctx->stream, ctx->partials, ctx->sc);
  FSX_TRY(launch_check(ctx, "residual check"));
  FSX_TRY(read_scalars(ctx));
  *stop = ctx->h_sc->stop;
  *rr_hi = ctx->h_sc->rr_hi;
  return FSX_OK;
}

int launch_fold_residual(fsx_ctx* ctx, double* x, const double* b, int* stop, double* rr_hi) {
  return residual_check(ctx, x, b, true, stop, rr_hi);
}
int launch_residual_only(fsx_ctx* ctx, double* x, const double* b, int* stop, double* rr_hi) {
  return residual_check(ctx, x, b, false, stop, rr_hi);
}

int solve_learned(fsx_ctx* ctx, const double* b, double* x, int* iters, double* relres, int truncate_at);
int solve_learned_tiled(fsx_ctx* ctx, const double* b, double* x, bool warm, int* iters, double* relres,
                        int truncate_at);

// |b|^2 (slot 0) and |r|^2 (slot 1)
__global__ __launch_bounds__(kBlock) void k_sumsq2(int n, const double* __restrict__ b, const double* __restrict__ r,
                                                   double* __restrict__ partials) {
  double v[2] = {0.0, 0.0};
#pragma unroll 4   // independent elements: their loads in flight together
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += kRedGrid * kBlock) {
    v[0] += b[i] * b[i];
    v[1] += r[i] * r[i];
  }
  write_partials<2>(v, partials);
}

// rh = b - Z x in FP64 and the partial sums of |b|^2 (slot 0) and |rh|^2
// (slot 1) of ctx->partials, with no read-back (the tiled learned solver's
// warm start takes them on the device)
int launch_warm_residual(fsx_ctx* ctx, const double* x, const double* b, double* rh) {
  if (ctx->warm_r != rh) launch_zresidual(ctx, x, b, rh);   // else the projection start formed it
  ctx->warm_r = nullptr;
  hipLaunchKernelGGL(k_sumsq2, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, ctx->NNp, b, rh, ctx->partials);
  return launch_check(ctx, "warm residual");
}

// Cholesky of Z (the MATLAB prototype's pressure step: factor once,
// ML:1876-1885, then two triangular solves, ML:2021-2030), then the FP64
// residual |b - Z x| / |b|, which must meet rtol like the iterative solvers'
static int solve_direct(fsx_ctx* ctx, const double* b, double* x, int* iters, double* relres) {
  const int n = ctx->NNp;
  if (ctx->direct_order == FSX_DIRECT_ND) {
    FSX_TRY(mf_factor(ctx));
    FSX_TRY(mf_apply(ctx, b, x));
  } else {
    FSX_TRY(direct_factor(ctx));
    FSX_TRY(direct_apply(ctx, b, x));
  }
  launch_zresidual(ctx, x, b, ctx->w0r);
  hipLaunchKernelGGL(k_sumsq2, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, n, b, ctx->w0r, ctx->partials);
  FSX_TRY(launch_check(ctx, "k_sumsq2"));
  double h[2 * kRedGrid];
  FSX_HIP(hipMemcpyAsync(h, ctx->partials, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
  FSX_HIP(hipStreamSynchronize(ctx->stream));
  double bb = 0.0, rr = 0.0;
  for (int i = 0; i < kRedGrid; i++) { bb += h[i]; rr += h[kRedGrid + i]; }
  *iters = 0;
  *relres = bb > 0 ? std::sqrt(rr / bb) : 0.0;
  if (!(rr == rr)) return fail(ctx, FSX_ENUMERIC, "direct solve residual is NaN/Inf");
  if (*relres > ctx->rtol) return fail(ctx, FSX_ENOCONV, "direct solve: residual above rtol");
  return FSX_OK;
}

// the learned solver in its tiled two-level form (a structured pressure grid
// whose Z rows compress into classes, and the multiplicative form selected)
static bool learned_tiled(const fsx_ctx* ctx) {
  return ctx->solver_mode == FSX_SOLVER_LEARNED && ctx->unet && ctx->tile.ntiles > 0 && ctx->learned_form == 1;
}

// the selected solver from x = 0
static int solve_cold(fsx_ctx* ctx, const double* b, double* x, int* iters, double* relres) {
  if (ctx->solver_mode == FSX_SOLVER_DIRECT) return solve_direct(ctx, b, x, iters, relres);
  if (learned_tiled(ctx)) return solve_learned_tiled(ctx, b, x, false, iters, relres, 0);
  if (ctx->solver_mode == FSX_SOLVER_LEARNED && ctx->unet) return solve_learned(ctx, b, x, iters, relres, 0);
  if (ctx->solver_mode == FSX_SOLVER_MIXED_IR || ctx->solver_mode == FSX_SOLVER_LEARNED)
    return solve_mixed(ctx, b, x, iters, relres);
  return solve_fp64(ctx, b, x, iters, relres);
}

// Warm start from the solution already in x (the reference's GPU path hands
// CUSP the previous Pdot, CU:425-442, which is zeroed only once, CU:268):
// r0 = b - Z x in FP64, then the selected solver on Z e = r0 from e = 0 with
// its relative tolerance scaled by |b| / |r0|, so that it stops at
// |r0 - Z e| <= rtol |b|, the cold solve's test; x += e. An x already
// within the tolerance takes no iteration.
static int solve_warm(fsx_ctx* ctx, const double* b, double* x, int* iters, double* relres) {
  const int n = ctx->NNp;
  if (ctx->warm_r != ctx->w0r) launch_zresidual(ctx, x, b, ctx->w0r);   // else the projection start formed it
  ctx->warm_r = nullptr;
  hipLaunchKernelGGL(k_sumsq2, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, n, b, ctx->w0r, ctx->partials);
  FSX_TRY(launch_check(ctx, "k_sumsq2"));
  double h[2 * kRedGrid];
  FSX_HIP(hipMemcpyAsync(h, ctx->partials, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
  FSX_HIP(hipStreamSynchronize(ctx->stream));
  double bb = 0.0, rr = 0.0;
  for (int i = 0; i < kRedGrid; i++) { bb += h[i]; rr += h[kRedGrid + i]; }
  if (!(rr == rr)) return fail(ctx, FSX_ENUMERIC, "warm start residual is NaN/Inf");
  if (rr <= ctx->rtol * ctx->rtol * bb) {
    *iters = 0;
    *relres = bb > 0 ? std::sqrt(rr / bb) : 0.0;
    return FSX_OK;
  }
  const double saved = ctx->rtol;
  ctx->rtol = std::min(1.0, saved * std::sqrt(bb / rr));
  double rel_e = 0.0;
  const int rc = solve_cold(ctx, ctx->w0r, ctx->w0e, iters, &rel_e);
  ctx->rtol = saved;
  if (rc != FSX_OK && rc != FSX_ENOCONV) return rc;
  FSX_TRY(launch_add(ctx, x, ctx->w0e, n));
  *relres = bb > 0 ? rel_e * std::sqrt(rr / bb) : 0.0;
  return rc;
}

// ------------------------------------------------------------ projection initial guess
// FSX_X0_PROJECTION (P. F. Fischer, "Projection techniques for iterative
// solution of Ax = b with successive right-hand sides", CMAME 163, 1998):
// x0 = X a, the Z-norm best approximation of Z^-1 b in the span of the last
// kProjMax solutions X, from (X^T Z X) a = X^T b. Z x_j is kept beside x_j
// (proj_b), from the solve's own final FP64 residual, so G[i][j] =
// x_i . (Z x_j) needs no product with Z; the Gram system is solved on the
// host by a Cholesky factorization that drops the vectors adding under
// 1e-10 of their own Z norm to the span (newest first).
// 16 kept solutions (kProjMax): 64^3 cavity, learned solve 48.4 ms over 20
// steps (8: 53.2 ms with Z x_j taken as b_j; 24: 51.1 ms, the longer dot and
// combination passes outweighing the iterations saved)
struct ProjSum {
  const double* x[kProjMax];
  const double* z[kProjMax];   // Z x of each
  double c[kProjMax];
  int n;
};

// X^T b and, when a slot s was recorded since the last guess (s >= 0), its
// row and column of G: block (chunk, j) sums its contiguous chunk of x_j . b
// (part row j), x_s . Z x_j (m + j) and x_j . Z x_s (2m + j). (One block a
// chunk for all j, 3m sums a thread, took 123 us a guess at 64^3.)
__global__ __launch_bounds__(kBlock) void k_proj_dots1(int n, int m, int s, const double* __restrict__ X,
                                                       const double* __restrict__ Bv, const double* __restrict__ b,
                                                       double* __restrict__ part) {
  __shared__ double red[3 * (kBlock / kWave)];
  const int j = blockIdx.y;
  const int per = (n + kRedGrid - 1) / kRedGrid;
  const int lo = blockIdx.x * per, hi = lo + per < n ? lo + per : n;
  const double* __restrict__ xj = X + (size_t)j * n;
  const double* __restrict__ zj = Bv + (size_t)j * n;
  const double* __restrict__ xs = X + (size_t)(s < 0 ? 0 : s) * n;
  const double* __restrict__ zs = Bv + (size_t)(s < 0 ? 0 : s) * n;
  double v[3] = {0.0, 0.0, 0.0};
  for (int i = lo + threadIdx.x; i < hi; i += kBlock) {
    const double x = xj[i];
    v[0] += x * b[i];
    if (s >= 0) {
      v[1] += xs[i] * zj[i];
      v[2] += x * zs[i];
    }
  }
  block_sum_all<3>(v, red);
  if (threadIdx.x == 0) {
    part[j * kRedGrid + blockIdx.x] = v[0];
    if (s >= 0) {
      part[(m + j) * kRedGrid + blockIdx.x] = v[1];
      part[(2 * m + j) * kRedGrid + blockIdx.x] = v[2];
    }
  }
}

// The nq dot products from k_proj_dots1's partials (a wavefront per dot
// product: each lane sums a strided quarter of its kRedGrid partials, then
// the lanes fold by shuffles) into the pinned publication, then its
// sequence number: the host waits on it instead of a copy and a stream sync.
// (One thread per dot product summing its partials in a chain took 10 us.)
constexpr int kPubBlock = 1024;
__global__ __launch_bounds__(kPubBlock) void k_proj_publish(int nq, const double* __restrict__ part, Published* pub,
                                                            unsigned seq) {
  const int lane = threadIdx.x & (kWave - 1);
  for (int q = threadIdx.x / kWave; q < nq; q += kPubBlock / kWave) {
    const double* p = part + q * kRedGrid;
    double t = 0.0;
#pragma unroll
    for (int blk = lane; blk < kRedGrid; blk += kWave) t += p[blk];
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) t += __shfl_down(t, o, kWave);
    if (lane == 0) reinterpret_cast<volatile double*>(pub->proj)[q] = t;
  }
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0) *reinterpret_cast<volatile unsigned*>(&pub->seq) = seq;
}

// x = sum_j c_j x_j and r = b - sum_j c_j Z x_j (= b - Z x to rounding), in
// the order of j
__global__ void k_proj_combine(int n, ProjSum S, const double* __restrict__ b, double* __restrict__ x,
                               double* __restrict__ r) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double v = 0.0, z = 0.0;
    for (int j = 0; j < S.n; j++) {
      v += S.c[j] * S.x[j][i];
      z += S.c[j] * S.z[j][i];
    }
    x[i] = v;
    r[i] = b[i] - z;
  }
}

void proj_reset(fsx_ctx* ctx) {
  ctx->proj_n = 0;
  ctx->proj_next = 0;
  ctx->proj_pending = 0;
}

static int proj_alloc(fsx_ctx* ctx) {
  if (ctx->proj_x) return FSX_OK;
  const size_t N = ctx->NNp;
  FSX_TRY(alloc(ctx, &ctx->proj_x, kProjMax * N));
  FSX_TRY(alloc(ctx, &ctx->proj_b, kProjMax * N));
  FSX_TRY(alloc(ctx, &ctx->proj_part, (size_t)kProjDots * kRedGrid));
  ctx->proj_G.assign(kProjMax * kProjMax, 0.0);
  proj_reset(ctx);
  return FSX_OK;
}

// x0 into x (left as it is while the window is empty), and r = b - Z x0
// into r, then ctx->warm_r = r (the warm start then skips that product);
// proj_pending is the bit mask of the slot recorded since the last guess (at
// most one: each record follows a guess, and a guess that fails empties the
// window)
static int proj_guess(fsx_ctx* ctx, const double* b, double* x, double* r) {
  FSX_TRY(proj_alloc(ctx));
  const int m = ctx->proj_n;
  if (m == 0) return FSX_OK;
  const size_t N = ctx->NNp;
  const unsigned stale = (unsigned)ctx->proj_pending;
  const int one = stale ? __builtin_ctz(stale) : -1;
  int k = 0;
  int gi[kProjDots], gj[kProjDots];
  for (int j = 0; j < m; j++) { gi[k] = -1; gj[k] = j; k++; }
  if (one >= 0) {
    for (int j = 0; j < m; j++) { gi[k] = one; gj[k] = j; k++; }
    for (int j = 0; j < m; j++) { gi[k] = j; gj[k] = one; k++; }
  }
  hipLaunchKernelGGL(k_proj_dots1, dim3(kRedGrid, m), dim3(kBlock), 0, ctx->stream, (int)N, m, one, ctx->proj_x,
                     ctx->proj_b, b, ctx->proj_part);
  const unsigned seq = ++ctx->pub_seq;
  hipLaunchKernelGGL(k_proj_publish, dim3(1), dim3(kPubBlock), 0, ctx->stream, k, ctx->proj_part, ctx->d_pub, seq);
  FSX_TRY(launch_check(ctx, "k_proj_dots"));
  FSX_TRY(wait_published(ctx, seq, "projection start"));
  double d[kProjMax] = {0.0};
  std::vector<double>& G = ctx->proj_G;
  for (int q = 0; q < k; q++) {
    const double t = const_cast<const double*>(ctx->h_pub->proj)[q];
    if (!std::isfinite(t)) {   // the guess is skipped (the solve meets its own test) and the window emptied
      proj_reset(ctx);
      return FSX_OK;
    }
    if (gi[q] < 0) d[gj[q]] = t;
    else G[gi[q] * kProjMax + gj[q]] = t;
  }
  ctx->proj_pending = 0;
  // Cholesky of the symmetrised Gram matrix over the kept slots, newest first
  int sel[kProjMax], ns = 0;
  double L[kProjMax][kProjMax] = {};
  for (int r = 0; r < m; r++) {
    const int c = (ctx->proj_next - 1 - r + 2 * kProjMax) % kProjMax;
    const double gcc = G[c * kProjMax + c];
    if (!(gcc > 0.0)) continue;
    double l[kProjMax], piv = gcc;
    for (int q = 0; q < ns; q++) {
      const int o = sel[q];
      double v = 0.5 * (G[c * kProjMax + o] + G[o * kProjMax + c]);
      for (int t = 0; t < q; t++) v -= L[q][t] * l[t];
      l[q] = v / L[q][q];
      piv -= l[q] * l[q];
    }
    if (!(piv > 1e-10 * gcc)) continue;
    for (int q = 0; q < ns; q++) L[ns][q] = l[q];
    L[ns][ns] = std::sqrt(piv);
    sel[ns++] = c;
  }
  if (ns == 0) return FSX_OK;
  double y[kProjMax], a[kProjMax];
  for (int q = 0; q < ns; q++) {
    double v = d[sel[q]];
    for (int t = 0; t < q; t++) v -= L[q][t] * y[t];
    y[q] = v / L[q][q];
  }
  for (int q = ns - 1; q >= 0; q--) {
    double v = y[q];
    for (int t = q + 1; t < ns; t++) v -= L[t][q] * a[t];
    a[q] = v / L[q][q];
  }
  ProjSum S{};
  S.n = ns;
  for (int q = 0; q < ns; q++) {
    S.x[q] = ctx->proj_x + sel[q] * N;
    S.z[q] = ctx->proj_b + sel[q] * N;
    S.c[q] = a[q];
  }
  hipLaunchKernelGGL(k_proj_combine, dim3(ew_grid((long long)N)), dim3(kBlock), 0, ctx->stream, (int)N, S, b, x, r);
  FSX_TRY(launch_check(ctx, "k_proj_combine"));
  ctx->warm_r = r;
  return FSX_OK;
}

// one window slot: xs = x and zs = b - w (w = 0: zs = b)
__global__ void k_proj_store(int n, const double* __restrict__ x, const double* __restrict__ b,
                             const double* __restrict__ w, double* __restrict__ xs, double* __restrict__ zs) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    xs[i] = x[i];
    zs[i] = w ? b[i] - w[i] : b[i];
  }
}

// the solution x of b into the window (the oldest slot when it is full),
// with Z x: b less the FP64 residual the mixed and learned solves leave in
// wh for their final x; the FP64 solver's recurrence residual is not kept,
// so there Z x is taken as b
static int proj_record(fsx_ctx* ctx, const double* b, const double* x) {
  FSX_TRY(proj_alloc(ctx));
  const size_t N = ctx->NNp;
  const int s = ctx->proj_next;
  const bool resid = ctx->solver_mode == FSX_SOLVER_MIXED_IR || ctx->solver_mode == FSX_SOLVER_LEARNED;
  hipLaunchKernelGGL(k_proj_store, dim3(ew_grid((long long)N)), dim3(kBlock), 0, ctx->stream, (int)N, x, b,
                     resid ? (const double*)ctx->wh : nullptr, ctx->proj_x + s * N, ctx->proj_b + s * N);
  FSX_TRY(launch_check(ctx, "k_proj_store"));
  ctx->proj_next = (s + 1) % kProjMax;
  ctx->proj_n = std::min(kProjMax, ctx->proj_n + 1);
  ctx->proj_pending = 1 << s;
  return FSX_OK;
}

int pressure_solve(fsx_ctx* ctx, const double* b, double* x, int* iters, double* relres) {
  int rc;
  timing_begin(ctx);
  // (a direct solve does not use an initial guess; the tiled learned solver
  // starts from x on the device)
  const bool direct = ctx->solver_mode == FSX_SOLVER_DIRECT;
  const bool proj = ctx->x0_policy == FSX_X0_PROJECTION && !direct;
  const bool warm = (ctx->x0_policy == FSX_X0_PREVIOUS || proj) && !direct;
  // the projection start also forms b - Z x0 where the warm start reads it
  ctx->warm_r = nullptr;
  if (proj) FSX_TRY(proj_guess(ctx, b, x, learned_tiled(ctx) ? ctx->wh : ctx->w0r));
  if (warm && learned_tiled(ctx)) rc = solve_learned_tiled(ctx, b, x, true, iters, relres, 0);
  else if (warm) rc = solve_warm(ctx, b, x, iters, relres);
  else rc = solve_cold(ctx, b, x, iters, relres);
  ctx->warm_r = nullptr;
  if (rc == FSX_OK || rc == FSX_ENOCONV) timing_end(ctx, *iters);
  ctx->last_relres = *relres;
  if (proj && rc == FSX_OK) FSX_TRY(proj_record(ctx, b, x));
  if (ctx->capture_hook && rc == FSX_OK) ctx->capture_hook(ctx, b, x);
  return rc;
}

// ------------------------------------------------------------ SpMV timing
__global__ void k_arm(SolveScalars* sc) { sc->stop = 0; }

template <class T>
__global__ void k_fill_t(T* __restrict__ p, T v, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

template <class T>
static int bench_spmv_t(fsx_ctx* ctx, T* p, T* q, int reps, double* ms) {
  const int n = ctx->NNp;
  hipLaunchKernelGGL(k_fill_t<T>, dim3(1024), dim3(kBlock), 0, ctx->stream, p, (T)1, n);
  hipLaunchKernelGGL(k_arm, dim3(1), dim3(1), 0, ctx->stream, ctx->sc);
  for (int w = 0; w < 3; w++)
    launch_zspmv<T>(ctx, p, q, ctx->sc);
  FSX_HIP(hipEventRecord(ctx->ev0, ctx->stream));
  for (int r = 0; r < reps; r++)
    launch_zspmv<T>(ctx, p, q, ctx->sc);
  FSX_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  FSX_HIP(hipEventSynchronize(ctx->ev1));
  FSX_TRY(launch_check(ctx, "k_pcg_spmv (timing)"));
  float t = 0;
  FSX_HIP(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
  *ms = t / reps;
  return FSX_OK;
}

}  // namespace fsx

extern "C" int fsx_bench_spmv(fsx_ctx* ctx, int precision, int reps, double* ms_per_launch, double* bytes_per_launch) {
  using namespace fsx;
  if (!ctx || reps < 1 || !ms_per_launch) return FSX_EINVAL;
  if (!ctx->z_built) return fail(ctx, FSX_ESTATE, "build Z first");
  double ms = 0;
  const double n = ctx->NNp;
  double bytes;
  // algorithmic bytes: values and columns as the solvers store them, p read
  // once, q written once.
  //  FP64: values of the rows without structural zeros + column indices (or,
  //  with the row-pattern dictionary, one pattern start per row plus the
  //  dictionary) + row pointers.
  //  FP32 row classes: the launch plan (16 bytes per chunk of rows, 4 per
  //  row) + the class table; otherwise
  //  the FP32 copy on the rows of the FP64 form.
  if (precision == 0) {
    FSX_TRY(bench_spmv_t<double>(ctx, ctx->wp, ctx->wq, reps, &ms));
    const double nnz = (double)ctx->nnzZr;
    const double cols = ctx->Zr_pat ? 4.0 * n + 4.0 * (ctx->Zr_patlen + ctx->Zr_npat + 1) : 4.0 * nnz;
    bytes = 8.0 * nnz + cols + 4.0 * (n + 1) + 16.0 * n;
  } else {
    FSX_TRY(bench_spmv_t<float>(ctx, ctx->fp, ctx->fq, reps, &ms));
    if (ctx->Zc_chunk) {
      bytes = 16.0 * ctx->Zc_nchunk + 4.0 * n + 8.0 * ctx->Zc_len + 8.0 * n;
    } else {
      const double nnz = (double)ctx->nnzZr;
      const double cols = ctx->Zr_pat ? 4.0 * n + 4.0 * (ctx->Zr_patlen + ctx->Zr_npat + 1) : 4.0 * nnz;
      bytes = 4.0 * nnz + cols + 4.0 * (n + 1) + 8.0 * n;
    }
  }
  *ms_per_launch = ms;
  if (bytes_per_launch) *bytes_per_launch = bytes;
  return FSX_OK;
}

// Times `reps` back-to-back launches of one row kernel of the time step on
// the context's stream (HIP events) and returns its algorithmic HBM bytes:
// every matrix entry, row pointer and vector element it touches counted once
// (int32 indices, FP64 values and vectors). The step vectors are clobbered.
extern "C" int fsx_bench_kernel(fsx_ctx* ctx, int kernel, int reps, double* ms_per_launch, double* bytes_per_launch) {
  using namespace fsx;
  if (!ctx || reps < 1 || !ms_per_launch) return FSX_EINVAL;
  if (!ctx->K || !ctx->G || !ctx->GT) return fail(ctx, FSX_ESTATE, "fsx_bench_kernel: assemble first");
  if (kernel < FSX_KERNEL_STEP1 || kernel > FSX_KERNEL_STEP1_NODES) return fail(ctx, FSX_EINVAL, "fsx_bench_kernel: kernel id");
  const double NN = ctx->NN, NNp = ctx->NNp, nK = (double)ctx->nnzM, nG = (double)ctx->nnzG, NE = ctx->NE;
  const bool elems = kernel == FSX_KERNEL_STEP1_ELEMENTS || kernel == FSX_KERNEL_STEP1_NODES;
  if (elems) {
    // the FP32 shadows the element form reads, as an inner iteration >= 2 has
    // them: UnpHalf_prev, Acc_prev and Un
    if (ctx->rec_mem) {
      hipLaunchKernelGGL(k_rec, dim3(ew_grid(ctx->NN)), dim3(kBlock), 0, ctx->stream, (long long)ctx->NN,
                         ctx->UnpHalf_prev, reinterpret_cast<float4*>(ctx->Urec_prev), 2);
      hipLaunchKernelGGL(k_rec, dim3(ew_grid(ctx->NN)), dim3(kBlock), 0, ctx->stream, (long long)ctx->NN, ctx->Acc_prev,
                         reinterpret_cast<float4*>(ctx->Arec_prev), 2);
      hipLaunchKernelGGL(k_rec, dim3(ew_grid(ctx->NN)), dim3(kBlock), 0, ctx->stream, (long long)ctx->NN, ctx->Un,
                         reinterpret_cast<float4*>(ctx->Ucmp), 1);
      FSX_TRY(launch_check(ctx, "k_rec"));
      ctx->urec_prev_ok = ctx->arec_prev_ok = ctx->ucmp_un = true;
    }
    ctx->kacc_pending = true;
    if (!step1_elements_ok(ctx))
      return fail(ctx, FSX_ESTATE, "fsx_bench_kernel: the element form of step1 needs FP32 matrices and convection, "
                                   "lumped mass and the assembled K");
  }
  const double vb = ctx->Kf ? 4.0 : 8.0;   // bytes per stored matrix value
  if (kernel == FSX_KERNEL_STEP1 && ctx->Kf && ctx->rec_mem && ctx->mass_mode == FSX_MASS_LUMPED) {
    // the FP32 shadows of UnpHalf_prev and Acc_prev, as the inner iterations have them
    hipLaunchKernelGGL(k_rec, dim3(ew_grid(ctx->NN)), dim3(kBlock), 0, ctx->stream, (long long)ctx->NN, ctx->UnpHalf_prev,
                       reinterpret_cast<float4*>(ctx->Urec_prev), 2);
    hipLaunchKernelGGL(k_rec, dim3(ew_grid(ctx->NN)), dim3(kBlock), 0, ctx->stream, (long long)ctx->NN, ctx->Acc_prev,
                       reinterpret_cast<float4*>(ctx->Arec_prev), 2);
    FSX_TRY(launch_check(ctx, "k_rec"));
    ctx->urec_prev_ok = ctx->arec_prev_ok = true;
  }
  // K's columns: a pattern start per row plus the dictionary, or a column per entry
  const double kcols = ctx->Mrowpat ? 4.0 * NN + 4.0 * (double)ctx->Mpatlen : 4.0 * nK;
  auto launch = [&]() {
    switch (kernel) {
      case FSX_KERNEL_STEP1: ctx->kacc_pending = true; ctx->gpn_valid = true; (void)launch_step1(ctx, 2); break;
      case FSX_KERNEL_KACC: (void)launch_kacc(ctx); break;
      case FSX_KERNEL_STEP2_RHS: (void)launch_step2_rhs(ctx, 2); break;
      case FSX_KERNEL_STEP1_ELEMENTS: (void)launch_elem_pass(ctx, true); break;
      case FSX_KERNEL_STEP1_NODES:
        ctx->gpn_valid = true;
        (void)launch_nodes_pass(ctx, true, true);
        ctx->urec_ok = ctx->drec_ok = ctx->drec_kacc = true;   // as step1 leaves them for step2
        break;
      default: launch_step3_rows(ctx, 2); break;
    }
  };
  double bytes = 0;
  switch (kernel) {
    case FSX_KERNEL_STEP1: {  // K (col + value), G Pn (kept: 3 doubles per row), Uprev, R1 r/w, Un, MdInv, bc,
                              // UnpHalf, Acc_prev, KAcc: the form of the second and later inner iterations;
                              // with the FP32 shadows Uprev and Acc_prev are 16 bytes per node and
                              // the shadows of UnpHalf and of step2's operand are written (12 + 12
                              // bytes; the latter reads MdOrigInv, 8)
      const bool wrec = ctx->Kf && ctx->rec_mem && ctx->mass_mode == FSX_MASS_LUMPED;
      const bool rrec = wrec && ctx->urec_prev_ok && ctx->arec_prev_ok;
      const double vin = rrec ? 16.0 : 24.0;
      bytes = vb * nK + kcols + 4.0 * (NN + 1) + 24.0 * NN + vin * NN + 48.0 * NN + 24.0 * NN + 8.0 * NN + 1.0 * NN +
              24.0 * NN + vin * NN + 24.0 * NN + (wrec ? 32.0 * NN : 0.0);
      break;
    }
    case FSX_KERNEL_KACC:    // K, Acc_prev, KAcc
      bytes = vb * nK + kcols + 4.0 * (NN + 1) + 24.0 * NN + 24.0 * NN;
      break;
    case FSX_KERNEL_STEP2_RHS: {  // G^T (col + 3 values), UnpHalf, MdOrigInv, KAcc, R2; or G^T, d's FP32 shadow,
                                  // R2. G^T as row classes: each row's class word and the class entries (16
                                  // bytes each, once), with the columns and row pointers, or each row's first
                                  // column where the classes carry the offsets
      const double gt = !g_classes_on(ctx) ? (4.0 + 3.0 * vb) * nG
                        : (ctx->GTc_base ? 8.0 * NNp : 4.0 * nG + 4.0 * NNp + 4.0 * (NNp + 1)) + 16.0 * (double)ctx->gtc_nent;
      const double rowp = g_classes_on(ctx) ? 0.0 : 4.0 * (NNp + 1);   // G^T's row pointers
      if (ctx->GTf && ctx->drec_ok && ctx->drec_kacc)
        bytes = gt + rowp + 16.0 * NN + 8.0 * NNp;
      else
        bytes = gt + rowp + 24.0 * NN + 8.0 * NN + 24.0 * NN + 8.0 * NNp;
      break;
    }
    case FSX_KERNEL_STEP1_ELEMENTS:   // per element its 27 node ids, corner offsets (24 floats) and rows
                                      // (27 x 6 floats); per node the FP32 records of Un (16 bytes) and of
                                      // UnpHalf_prev with Acc_prev (32)
      bytes = (27.0 * 4 + 24.0 * 4 + 27.0 * 6 * 4) * NE + 48.0 * NN;
      break;
    case FSX_KERNEL_STEP1_NODES:      // the incidence (pointer, 27 entries an element), the element rows;
                                      // G Pn, Un, MdInv, bc, MdOrigInv read; R1, UnpHalf, KAcc written, the
                                      // shadows of UnpHalf and of step2's operand (12 + 12 bytes)
      bytes = 4.0 * (NN + 1) + (27.0 * 4 + 27.0 * 6 * 4) * NE + (24.0 + 24.0 + 8.0 + 1.0 + 8.0) * NN +
              (24.0 + 24.0 + 24.0 + 12.0 + 12.0) * NN;
      break;
    default: {               // G, Pdot, KAcc, bc, MdInv, UnpHalf, Acc (+ its FP32 shadow), Unp1; with the
                             // element form of step1 also G Pn read and G Pnp1 written (the carry). G as
                             // row classes: each row's first column and class word (8 bytes) and the
                             // class entries (16 bytes each, once) in place of its columns and values
      const bool carry = ctx->Gf && ctx->gpn_valid && ctx->use_step1_elements && ctx->GPn_next;
      const double g = g_classes_on(ctx) ? 8.0 * NN + 16.0 * (double)ctx->gc_nent : (4.0 + 3.0 * vb) * nG + 4.0 * (NN + 1);
      bytes = g + 8.0 * NNp + 24.0 * NN + 1.0 * NN + 8.0 * NN + 24.0 * NN + 24.0 * NN +
              24.0 * NN + (ctx->Gf && ctx->rec_mem ? 12.0 * NN : 0.0) + (carry ? 48.0 * NN : 0.0);
      break;
    }
  }
  FSX_TRY(settle_deferred(ctx));
  if (kernel == FSX_KERNEL_STEP1_NODES) FSX_TRY(launch_elem_pass(ctx, true));   // the rows the node pass sums
  for (int w = 0; w < 2; w++) launch();
  FSX_HIP(hipEventRecord(ctx->ev0, ctx->stream));
  for (int r = 0; r < reps; r++) launch();
  FSX_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  FSX_HIP(hipEventSynchronize(ctx->ev1));
  FSX_TRY(launch_check(ctx, "fsx_bench_kernel"));
  float t = 0;
  FSX_HIP(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
  *ms_per_launch = t / reps;
  if (bytes_per_launch) *bytes_per_launch = bytes;
  return FSX_OK;
}

extern "C" int fsx_z_forms(fsx_ctx* ctx, long long* nnz_solver, int* nclasses, long long* class_entries) {
  if (!ctx) return FSX_EINVAL;
  if (!ctx->z_built) return fsx::fail(ctx, FSX_ESTATE, "build Z first");
  if (nnz_solver) *nnz_solver = ctx->nnzZr;
  if (nclasses) *nclasses = ctx->Zc_chunk ? ctx->Zc_n : 0;
  if (class_entries) *class_entries = ctx->Zc_chunk ? ctx->Zc_len : 0;
  return FSX_OK;
}

extern "C" int fsx_z_patterns(fsx_ctx* ctx, int* npatterns, long long* offsets) {
  if (!ctx) return FSX_EINVAL;
  if (npatterns) *npatterns = ctx->Zr_pat ? ctx->Zr_npat : 0;
  if (offsets) *offsets = ctx->Zr_pat ? ctx->Zr_patlen : 0;
  return FSX_OK;
}

// Times `reps` launches of a plain device copy of `bytes` (rounded down to
// whole doubles) between two scratch buffers: the reference figure for HBM
// bandwidth and the kernel of known traffic that calibrates the PMC byte
// counters. No reference counterpart.
extern "C" int fsx_bench_copy(fsx_ctx* ctx, size_t bytes, int reps, double* ms_per_launch) {
  using namespace fsx;
  if (!ctx || reps < 1 || bytes < 8 || !ms_per_launch) return FSX_EINVAL;
  const long long n = (long long)(bytes / 8);
  double *a = nullptr, *b = nullptr;
  FSX_TRY(alloc(ctx, &a, (size_t)n));
  int rc = alloc(ctx, &b, (size_t)n);
  if (rc != FSX_OK) { dev_free(ctx, a); return rc; }
  rc = launch_fill(ctx, a, 1.0, n);
  for (int w = 0; w < 2 && rc == FSX_OK; w++) rc = launch_copy(ctx, b, a, n);
  float t = 0;
  if (rc == FSX_OK) {
    (void)hipEventRecord(ctx->ev0, ctx->stream);
    for (int r = 0; r < reps && rc == FSX_OK; r++) rc = launch_copy(ctx, b, a, n);
    (void)hipEventRecord(ctx->ev1, ctx->stream);
    if (hipEventSynchronize(ctx->ev1) != hipSuccess) rc = fail(ctx, FSX_EHIP, "fsx_bench_copy: event sync");
    (void)hipEventElapsedTime(&t, ctx->ev0, ctx->ev1);
  }
  (void)hipStreamSynchronize(ctx->stream);
  dev_free(ctx, a);
  dev_free(ctx, b);
  *ms_per_launch = t / reps;
  return rc;
}

// ------------------------------------------------------------ allocation-free pressure solve
// fsx_solve_pressure_noalloc runs the FP64 Jacobi-PCG of fsx_solve_pressure
// with everything beyond the caller's b, x and the context's Z held in one
// file-scope device array: the solve scalars, the reduction partials and the
// Krylov vectors r, z, p, q. Each device that loads the library holds its own
// copy of the array until the process ends; its device address is taken once
// per device by scratch_setup (at fsx_create, with the device current), and
// only that address is ever handed to kernels (the array's name on the host
// is a stand-in, not a device address).
namespace fsx {

constexpr long long kNoallocMaxRows = 1LL << 20;              // pressure nodes (NNp) the scratch holds
constexpr size_t kScratchHead = 256;                          // SolveScalars, padded
constexpr size_t kScratchPartials = 8 * (size_t)kRedGrid;     // doubles of reduction partials
constexpr size_t kScratchBytes = kScratchHead + 8 * kScratchPartials + 4 * 8 * (size_t)kNoallocMaxRows;
static_assert(sizeof(SolveScalars) <= kScratchHead, "solve scalars fit the scratch head");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

int scratch_setup(fsx_ctx* ctx) {
  if (ctx->device < 0 || ctx->device >= kMaxDevices) return fail(ctx, FSX_EINVAL, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[ctx->device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(ctx, FSX_EHIP, std::string("hipGetSymbolAddress(scratch) -> ") + hipGetErrorString(e));
    g_scratch_ptr[ctx->device] = p;
  }
  ctx->scratch = g_scratch_ptr[ctx->device];
  return FSX_OK;
}

}  // namespace fsx

extern "C" int fsx_solve_pressure_noalloc(fsx_ctx* ctx, const double* b, double* x, int* iters, double* relres) {
  using namespace fsx;
  if (!ctx || !b || !x) return FSX_EINVAL;
  if (!ctx->z_built) return fail(ctx, FSX_ESTATE, "build Z first");
  if (!ctx->scratch) return fail(ctx, FSX_ESTATE, "scratch address not set up for this device");
  const int n = ctx->NNp;
  if ((long long)n > kNoallocMaxRows)
    return fail(ctx, FSX_EINVAL, "fsx_solve_pressure_noalloc: NNp " + std::to_string(n) + " exceeds the scratch maximum of " +
                                     std::to_string(kNoallocMaxRows) + " pressure nodes");
  unsigned char* base = static_cast<unsigned char*>(ctx->scratch);
  SolveScalars* sc = reinterpret_cast<SolveScalars*>(base);
  double* partials = reinterpret_cast<double*>(base + kScratchHead);
  double* r = partials + kScratchPartials;
  double* z = r + n;
  double* p = z + n;
  double* q = p + n;
  // k_pcg_init writes x, r, z, p and the partials, k_pcg_init_finish every
  // scalar the iterations read: nothing of a previous call is read
  hipLaunchKernelGGL(k_pcg_init<double>, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, n, b, ctx->Zdinv, x, r, z, p,
                     partials);
  hipLaunchKernelGGL(k_pcg_init_finish, dim3(1), dim3(kBlock), 0, ctx->stream, partials, sc, ctx->rtol, ctx->maxit);
  FSX_TRY(launch_check(ctx, "k_pcg_init (noalloc)"));
  int j = 0;
  int chunk = segment_guess(std::min(ctx->last_iters, ctx->maxit), 1.0, ctx->rtol, 4);
  for (;;) {
    for (int c = 0; c < chunk; c++)
      pcg_launch_iteration<double>(ctx, j + c, ctx->Zdinv, x, r, z, p, q, 0, 0.0, sc, partials);
    j += chunk;
    FSX_TRY(launch_check(ctx, "PCG iterations (noalloc)"));
    FSX_TRY(read_scalars(ctx, sc));
    if (ctx->h_sc->stop) break;
    if (j > ctx->maxit + 64) return fail(ctx, FSX_ESTATE, "PCG did not terminate");
    chunk = 4;
  }
  const SolveScalars& s = *ctx->h_sc;
  if (iters) *iters = s.iters;
  if (relres) *relres = s.bb > 0 ? std::sqrt(s.rr / s.bb) : 0.0;
  if (s.stop == 3) return fail(ctx, FSX_ENUMERIC, "PCG met NaN/Inf");
  if (s.stop == 2) { ctx->err = "PCG reached maxit"; return FSX_ENOCONV; }
  return FSX_OK;
}
