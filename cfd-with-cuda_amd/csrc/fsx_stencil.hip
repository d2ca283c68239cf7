// fsx_stencil.hip -- the pressure operator Z on a structured pressure grid,
// staged through LDS tiles, with the vector work of the learned solver's
// flexible-CG iteration fused into the three Z products it needs.
//
// On the structured meshes the pressure nodes form an nx*ny*nz grid (x
// fastest) and Z = sum_d G_d^T M^-1 G_d couples a node only to nodes at most
// 2 apart per axis (two hex8 elements away). A tile is every x of a few rows
// in y and z, at most 1024 rows; its nodes plus a halo of 2 are loaded into
// LDS once, combined on the way in (z + beta p, or the Jacobi-smoothed
// residual), and every row of the tile then sums its FP32 row class's entries
// from LDS, one row a thread. A tile's rows are sorted by class, so most
// wavefronts hold rows of one class and read its entries through scalar
// loads; the few whose rows mix classes (next to the domain boundary) read
// them per lane from the tile's entry block, staged in LDS beside the tile.
// Every tile is one pass of at most 1024 rows, whatever its position, and
// every global load of a launch is issued before the first wait.
#include "fsx_device.h"
#include "fsx_ztensor.h"

#include <algorithm>
#include <cstring>
#include <utility>
#include <vector>

namespace fsx {

namespace {

constexpr int kTileThreads = 1024;      // one row a thread
constexpr int kTileHalo = 2;            // Z couples nodes up to 2 apart per axis
constexpr int kTileMaxLds = 60 * 1024;  // bytes of LDS a tile may take (two tiles a CU)
constexpr int kTileSlotsMax = 17;       // tile nodes a thread loads in one pass (up to 17 x 1024)

}  // namespace

// Builds the tile plan from the FP32 row classes (row -> class, class entry
// pointers, column offsets and value bits, as build_row_dict returns them).
// Returns FSX_OK with ctx->tile.ntiles = 0 when the grid is not structured,
// a tile does not fit LDS, or some entry is not a radius-2 grid neighbour of
// its row (then the solvers keep the row-class form).
int build_tile_plan(fsx_ctx* ctx, const std::vector<int>& cls, const std::vector<int>& cp, const std::vector<int>& co,
                    const std::vector<unsigned>& cv) {
  for (void*& p : ctx->tile_mem)
    if (p) { dev_free(ctx, p); p = nullptr; }
  if (ctx->tile_part) { dev_free(ctx, ctx->tile_part); ctx->tile_part = nullptr; }
  ctx->tile = TilePlan{};
  const int nx = ctx->pg_nx, ny = ctx->pg_ny, nz = ctx->pg_nz;
  const int n = ctx->NNp;
  if (nx < 5 || ny < 5 || nz < 5 || nx > kTileThreads || (long long)nx * ny * nz != n || cls.size() != (size_t)n)
    return FSX_OK;
  const int nxy = nx * ny;
  // each class entry's grid displacement (unique for nx, ny >= 5)
  const size_t nent = co.size();
  std::vector<int> ddx(nent), ddy(nent), ddz(nent);
  for (size_t e = 0; e < nent; e++) {
    bool found = false;
    for (int dz = -2; dz <= 2 && !found; dz++)
      for (int dy = -2; dy <= 2 && !found; dy++) {
        const int dx = co[e] - dy * nx - dz * nxy;
        if (dx >= -2 && dx <= 2) { ddx[e] = dx; ddy[e] = dy; ddz[e] = dz; found = true; }
      }
    if (!found) return FSX_OK;
  }
  // every row's entries stay inside the grid (no wrap to the next line)
  for (int r = 0; r < n; r++) {
    const int i = r % nx, j = (r / nx) % ny, k = r / nxy;
    const int c = cls[r];
    if (cp[c + 1] - cp[c] > 255) return FSX_OK;
    for (int e = cp[c]; e < cp[c + 1]; e++) {
      const int a = i + ddx[e], b = j + ddy[e], d = k + ddz[e];
      if (a < 0 || a >= nx || b < 0 || b >= ny || d < 0 || d >= nz) return FSX_OK;
    }
  }
  // Tile shape: the most lines (ty x tz, ty >= tz) with nx * ty * tz <= 1024
  // rows whose tile, own-row operands and entry block fit LDS (small grids
  // put many classes in one tile, so they take smaller tiles).
  std::vector<std::pair<int, int>> shapes;
  for (int a = 1; a <= std::min(ny, kTileThreads / nx); a++)
    for (int c = 1; c <= std::min(a, nz); c++)
      if ((long long)nx * a * c <= kTileThreads) shapes.push_back({a, c});
  std::stable_sort(shapes.begin(), shapes.end(), [](const std::pair<int, int>& x, const std::pair<int, int>& y) {
    return x.first * x.second > y.first * y.second;
  });
  int ty = 0, tz = 0, lx = 0, ly = 0, lz = 0, nown = 0, tiles_y = 0, ntiles = 0;
  size_t estride = 0;
  std::vector<int2> ent(nent);
  std::vector<std::vector<int>> trows, tstart;
  std::vector<std::vector<int2>> tblk;
  for (const auto& sh : shapes) {
    ty = sh.first; tz = sh.second;
    lx = nx + 2 * kTileHalo; ly = ty + 2 * kTileHalo; lz = tz + 2 * kTileHalo;
    nown = nx * ty * tz;
    if ((long long)lx * ly * lz > kTileSlotsMax * (long long)kTileThreads || (long long)lx * ly * lz >= 65536) continue;
    for (size_t e = 0; e < nent; e++) ent[e] = make_int2(ddx[e] + lx * (ddy[e] + ly * ddz[e]), (int)cv[e]);
    tiles_y = (ny + ty - 1) / ty;
    ntiles = tiles_y * ((nz + tz - 1) / tz);
    // per tile: its rows sorted by class, and the entries of its classes
    trows.assign(ntiles, {});
    tblk.assign(ntiles, {});
    tstart.assign(ntiles, {});
    estride = 1;
    for (int t = 0; t < ntiles; t++) {
      const int ty0 = (t % tiles_y) * ty, tz0 = (t / tiles_y) * tz;
      std::vector<int>& tr = trows[t];
      for (int k = tz0; k < std::min(nz, tz0 + tz); k++)
        for (int j = ty0; j < std::min(ny, ty0 + ty); j++)
          for (int i = 0; i < nx; i++) tr.push_back(i + nx * (j + ny * k));
      std::stable_sort(tr.begin(), tr.end(), [&](int x, int y) { return cls[x] < cls[y]; });
      std::vector<int2>& blk = tblk[t];
      tstart[t].resize(tr.size());
      int last = -1, at = 0;
      for (size_t q = 0; q < tr.size(); q++) {
        const int c = cls[tr[q]];
        if (c != last) {
          at = (int)blk.size();
          blk.insert(blk.end(), ent.begin() + cp[c], ent.begin() + cp[c + 1]);
          last = c;
        }
        tstart[t][q] = at;
      }
      estride = std::max(estride, blk.size());
    }
    estride = (estride + 63) / 64 * 64;
    if (((size_t)lx * ly * lz + 4 * (size_t)nown) * sizeof(float) + estride * sizeof(int2) <= (size_t)kTileMaxLds &&
        estride <= 4 * (size_t)kTileThreads)
      break;
    ntiles = 0;
  }
  if (ntiles == 0) return FSX_OK;
  const size_t rstride = kTileThreads;
  std::vector<int4> rows((size_t)ntiles * rstride, make_int4(0, 0, 0, 0));
  std::vector<int2> tent((size_t)ntiles * estride, make_int2(0, 0));
  for (int t = 0; t < ntiles; t++) {
    const int ty0 = (t % tiles_y) * ty, tz0 = (t / tiles_y) * tz;
    for (size_t s = 0; s < trows[t].size(); s++) {
      const int r = trows[t][s];
      const int i = r % nx, j = (r / nx) % ny, k = r / nxy;
      const int li = (i + kTileHalo) + lx * ((j - ty0 + kTileHalo) + ly * (k - tz0 + kTileHalo));
      const int own = i + nx * ((j - ty0) + ty * (k - tz0));
      const int c = cls[r];
      rows[(size_t)t * rstride + s] = make_int4(r, li | (own << 16), cp[c], (cp[c + 1] - cp[c]) | (tstart[t][s] << 8));
    }
    std::copy(tblk[t].begin(), tblk[t].end(), tent.begin() + (size_t)t * estride);
  }
  int4* d_rows = nullptr; int2* d_ent = nullptr; int2* d_tent = nullptr;
  FSX_TRY(alloc(ctx, &d_rows, rows.size()));
  ctx->tile_mem[0] = d_rows;
  FSX_TRY(alloc(ctx, &d_ent, std::max<size_t>(1, ent.size())));
  ctx->tile_mem[1] = d_ent;
  FSX_TRY(alloc(ctx, &d_tent, tent.size()));
  ctx->tile_mem[2] = d_tent;
  FSX_HIP(hipMemcpy(d_rows, rows.data(), rows.size() * sizeof(int4), hipMemcpyHostToDevice));
  FSX_HIP(hipMemcpy(d_ent, ent.data(), ent.size() * sizeof(int2), hipMemcpyHostToDevice));
  FSX_HIP(hipMemcpy(d_tent, tent.data(), tent.size() * sizeof(int2), hipMemcpyHostToDevice));
  const int npart = (ntiles + 511) / 512 * 512;   // whole passes of 8 loads per lane of a wavefront
  FSX_TRY(alloc(ctx, &ctx->tile_part, (size_t)kTileSlots * npart));
  FSX_HIP(hipMemset(ctx->tile_part, 0, (size_t)kTileSlots * npart * sizeof(double)));
  TilePlan& T = ctx->tile;
  T.nx = nx; T.ny = ny; T.nz = nz;
  T.ty = ty; T.tz = tz; T.lx = lx; T.ly = ly; T.lz = lz;
  T.tiles_y = tiles_y; T.ntiles = ntiles; T.npart = npart;
  T.rstride = (int)rstride; T.estride = (int)estride;
  T.rows = d_rows; T.ent = d_ent; T.tent = d_tent;
  return FSX_OK;
}

// LDS floats of a tile of the Kronecker form: the tile, the own rows' three
// operands, the z pass's two products over (lx, ly, tz), the y pass's two
// over (lx, ty, tz) and the tile's axis coefficients (10 a grid point)
static size_t tensor_lds_floats(int nx, int ty, int tz) {
  const size_t lx = nx + 2 * kTileHalo, ly = ty + 2 * kTileHalo, lz = tz + 2 * kTileHalo;
  return lx * ly * lz + 3 * (size_t)nx * ty * tz + 2 * lx * ly * tz + 2 * lx * ty * tz + 10 * (size_t)(nx + ty + tz);
}

int build_tile_plan_tensor(fsx_ctx* ctx, const ZTensor& zt) {
  for (void*& p : ctx->tile_mem)
    if (p) { dev_free(ctx, p); p = nullptr; }
  if (ctx->tile_part) { dev_free(ctx, ctx->tile_part); ctx->tile_part = nullptr; }
  ctx->tile = TilePlan{};
  const int nx = zt.n[0], ny = zt.n[1], nz = zt.n[2];
  if (nx < 5 || ny < 5 || nz < 5 || nx > kTileThreads || (long long)nx * ny * nz != ctx->NNp) return FSX_OK;
  // the most lines (ty x tz, ty >= tz) a tile of at most 1024 rows that fits LDS
  int ty = 0, tz = 0, best = 0;
  for (int a = 1; a <= std::min(ny, kTileThreads / nx); a++)
    for (int c = 1; c <= std::min(a, nz); c++) {
      const long long L = (long long)(nx + 2 * kTileHalo) * (a + 2 * kTileHalo) * (c + 2 * kTileHalo);
      if ((long long)nx * a * c > kTileThreads || L > kTileSlotsMax * (long long)kTileThreads || L >= 65536) continue;
      if (tensor_lds_floats(nx, a, c) * sizeof(float) > (size_t)kTileMaxLds) continue;
      if (10 * (nx + a + c) > kTileThreads) continue;   // one coefficient a thread
      if (a * c > best) { best = a * c; ty = a; tz = c; }
    }
  if (best == 0) return FSX_OK;
  const int lx = nx + 2 * kTileHalo, ly = ty + 2 * kTileHalo, lz = tz + 2 * kTileHalo;
  const int tiles_y = (ny + ty - 1) / ty, ntiles = tiles_y * ((nz + tz - 1) / tz);
  const size_t rstride = kTileThreads;
  // rows in grid order: (row, LDS index | own index << 16, 0, 1), 0 past the tile's rows
  std::vector<int4> rows((size_t)ntiles * rstride, make_int4(0, 0, 0, 0));
  for (int t = 0; t < ntiles; t++) {
    const int ty0 = (t % tiles_y) * ty, tz0 = (t / tiles_y) * tz;
    size_t q = 0;
    for (int k = tz0; k < std::min(nz, tz0 + tz); k++)
      for (int j = ty0; j < std::min(ny, ty0 + ty); j++)
        for (int i = 0; i < nx; i++, q++) {
          const int r = i + nx * (j + ny * k);
          const int li = (i + kTileHalo) + lx * ((j - ty0 + kTileHalo) + ly * (k - tz0 + kTileHalo));
          const int own = i + nx * ((j - ty0) + ty * (k - tz0));
          rows[(size_t)t * rstride + q] = make_int4(r, li | (own << 16), 0, 1);
        }
  }
  std::vector<float> coef(10 * (size_t)(nx + ny + nz));
  size_t at = 0;
  for (int d = 0; d < 3; d++)
    for (int i = 0; i < zt.n[d]; i++) {
      for (int c = 0; c < 5; c++) coef[at++] = (float)zt.k[d][5 * (size_t)i + c];
      for (int c = 0; c < 5; c++) coef[at++] = (float)zt.m[d][5 * (size_t)i + c];
    }
  int4* d_rows = nullptr; float* d_coef = nullptr;
  FSX_TRY(alloc(ctx, &d_rows, rows.size()));
  ctx->tile_mem[0] = d_rows;
  FSX_TRY(alloc(ctx, &d_coef, coef.size()));
  ctx->tile_mem[1] = d_coef;
  FSX_HIP(hipMemcpy(d_rows, rows.data(), rows.size() * sizeof(int4), hipMemcpyHostToDevice));
  FSX_HIP(hipMemcpy(d_coef, coef.data(), coef.size() * sizeof(float), hipMemcpyHostToDevice));
  const int npart = (ntiles + 511) / 512 * 512;   // whole passes of 8 loads per lane of a wavefront
  FSX_TRY(alloc(ctx, &ctx->tile_part, (size_t)kTileSlots * npart));
  FSX_HIP(hipMemset(ctx->tile_part, 0, (size_t)kTileSlots * npart * sizeof(double)));
  TilePlan& T = ctx->tile;
  T.nx = nx; T.ny = ny; T.nz = nz;
  T.ty = ty; T.tz = tz; T.lx = lx; T.ly = ly; T.lz = lz;
  T.tiles_y = tiles_y; T.ntiles = ntiles; T.npart = npart;
  T.rstride = (int)rstride; T.estride = 0;
  T.rows = d_rows; T.ent = nullptr; T.tent = nullptr;
  T.tensor = 1; T.coef = d_coef; T.zp = zt.zero_p; T.pen = (float)zt.pen;
  return FSX_OK;
}

// ------------------------------------------------------------ tiled kernels
// One block of 1024 threads per tile, one row a thread. MODE:
//  kTileDir    (the iteration's first kernel) decides from the partial sums of
//              the previous iteration whether to stop, then p = z + beta p_in
//              (beta = -z.q / p.q of the last update, 0 after a start or a
//              replaced residual), q = Z p; slot 0 gets p.q.
//  kTileSmooth alpha = r.p / p.q; r' = r - alpha q, x += alpha p; the tile holds
//              t = omega dinv r' and r1 = r' - Z t (the network's input);
//              slots 1, 2 get r'.r', r'.p.
//  kTilePre    the same without the update (start, or a replaced residual):
//              r1 = r - Z (omega dinv r); slots 1, 2 get r.r, r.p_in.
//  kTilePost   the tile holds z2 = omega dinv r + C r1 (the network's output
//              added to the pre-smoothing); r2 = r - Z z2, z = z2 + omega dinv
//              r2; slots 3, 4 get r.z, z.q.
// The solve scalars change only in block 0 of kTileDir (decision, iteration
// count, r.p) and of kTileSmooth (p.q of the update); every other block reads
// values that an earlier launch wrote, and all blocks sum the same partials in
// the same order, so they reach the same decision. The plan's tables are
// __restrict__ parameters, so that wave-uniform reads of them become scalar
// loads; the operands the epilogue needs at a row come from LDS, where the
// tile load put them. At most 64 VGPRs (8 wavefronts a SIMD), so two tiles
// share a CU and every tile of the 65^3 grid is resident at once.
//
// TEN: Z as Kronecker products of 1D stencils (a tensor-product box grid):
// Z = Kx(x)My(x)Mz + Mx(x)Ky(x)Mz + Mx(x)My(x)Kz applied to the tile by
// passes along z (a = Mz t, b = Kz t over every x and y of the tile, own z),
// along y (c = My a, de = Ky a + My b over every x, own y and z) and, per
// row, along x (Kx c + Mx de): 45 multiply-adds a row instead of the 81 of
// its class list, with each pass's products staged in LDS.
struct TileGeom {
  int nx, ny, nz, ty, tz, lx, ly, lz, tiles_y, ntiles, npart, rstride, estride;
  const float* coef;
  int zp;
  float pen;
};

template <int MODE, int TS, bool TEN>
__global__ __launch_bounds__(kTileThreads, 8) void k_tile(TileGeom T, const int4* __restrict__ rows,
                                                          const int2* __restrict__ ent, const int2* __restrict__ tent,
                                                          TileVecs A) {
  extern __shared__ float tile[];
  SolveScalars* sc = A.sc;
  if (sc->stop) return;   // uniform
  const int b = blockIdx.x;
  const int L = T.lx * T.ly * T.lz;
  const int nown = T.nx * T.ty * T.tz;
  float* e0s = tile + L;          // own rows' operands: residual
  float* e1s = e0s + nown;        // direction, or dinv (kTilePost)
  float* e2s = e1s + nown;        // q (kTilePost), x (kTileSmooth)
  int2* tes = reinterpret_cast<int2*>(e2s + nown + ((L + 3 * nown) & 1));   // the tile's entry block (8-byte aligned)
  // the Kronecker form's products and coefficients (TEN)
  const int n1 = T.lx * T.ly * T.tz, n2 = T.lx * T.ty * T.tz;
  float* a1s = e2s + nown;
  float* b1s = a1s + n1;
  float* c2s = b1s + n1;
  float* d2s = c2s + n2;
  float* cxs = d2s + n2;          // nx x 10
  float* cys = cxs + 10 * T.nx;   // ty x 10
  float* czs = cys + 10 * T.ty;   // tz x 10
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), lane = threadIdx.x & (kWave - 1);
  // --- one pass of loads: this thread's row, the tile's entry block, the
  // tile nodes and the own rows' operands, then the partial sums
  const int4 rw = rows[(size_t)b * T.rstride + threadIdx.x];
  const int ty0 = (b % T.tiles_y) * T.ty, tz0 = (b / T.tiles_y) * T.tz;
  constexpr int NE = TEN ? 1 : 4;   // entries of the block a thread copies (estride <= 4 x 1024)
  int2 eb[NE];
  float cf = 0.f;                   // TEN: this thread's axis coefficient for the LDS tables
  const int ncf = 10 * (T.nx + T.ty + T.tz);
  if constexpr (TEN) {
    // coefficient q of the tile's tables: x rows, then the own y rows, then
    // the own z rows (zero past the grid)
    const int q = threadIdx.x;
    if (q < ncf) {
      const int row = q / 10, c = q - 10 * row;
      int g = -1;
      if (row < T.nx) g = row;
      else if (row < T.nx + T.ty) { const int j = ty0 + row - T.nx; g = j < T.ny ? T.nx + j : -1; }
      else { const int k = tz0 + row - T.nx - T.ty; g = k < T.nz ? T.nx + T.ny + k : -1; }
      cf = g >= 0 ? T.coef[10 * g + c] : 0.f;
    }
  } else {
#pragma unroll
    for (int u = 0; u < NE; u++) {
      const int q = threadIdx.x + u * kTileThreads;
      eb[u] = tent[(size_t)b * T.estride + (q < T.estride ? q : 0)];
    }
  }
  int g[TS], own[TS];
  bool in[TS];
#pragma unroll
  for (int u = 0; u < TS; u++) {
    const int e = threadIdx.x + u * kTileThreads;
    const unsigned ee = e < L ? (unsigned)e : 0u;
    const unsigned ex = ee % (unsigned)T.lx, rest = ee / (unsigned)T.lx;
    const unsigned ey = rest % (unsigned)T.ly, ez = rest / (unsigned)T.ly;
    const int gi = (int)ex - kTileHalo, gj = ty0 + (int)ey - kTileHalo, gk = tz0 + (int)ez - kTileHalo;
    in[u] = e < L && gi >= 0 && gi < T.nx && gj >= 0 && gj < T.ny && gk >= 0 && gk < T.nz;
    g[u] = in[u] ? gi + T.nx * (gj + T.ny * gk) : 0;
    const int oy = (int)ey - kTileHalo, oz = (int)ez - kTileHalo;
    own[u] = (in[u] && oy >= 0 && oy < T.ty && oz >= 0 && oz < T.tz) ? gi + T.nx * (oy + T.ty * oz) : -1;
  }
  float va[TS], vb[TS], vc[TS], vd[TS], ve[TS];
#pragma unroll
  for (int u = 0; u < TS; u++) {
    const int gg = g[u];
    if constexpr (MODE == kTileDir) { va[u] = A.z[gg]; vb[u] = A.pin[gg]; }
    else if constexpr (MODE == kTileSmooth) { va[u] = A.rin[gg]; vb[u] = A.q[gg]; vc[u] = A.dinv[gg]; }
    else if constexpr (MODE == kTilePre) { va[u] = A.rin[gg]; vc[u] = A.dinv[gg]; }
    else { va[u] = A.z2[gg]; }
  }
#pragma unroll
  for (int u = 0; u < TS; u++) {   // operands of the own rows only
    const int gg = own[u] >= 0 ? g[u] : 0;
    if constexpr (MODE == kTileSmooth) { vd[u] = A.pin[gg]; ve[u] = A.x[gg]; }
    else if constexpr (MODE == kTilePre) vd[u] = A.pin[gg];
    else if constexpr (MODE == kTilePost) { vb[u] = A.rin[gg]; vc[u] = A.dinv[gg]; vd[u] = A.q[gg]; }
  }
  // the partial sums last, so that their wait covers every load above
  // (a slot holds npart entries, a multiple of 512, zero past ntiles)
  constexpr int NSUM = MODE == kTileDir ? 4 : (MODE == kTileSmooth ? 1 : 0);
  constexpr int FIRST = MODE == kTileDir ? 1 : 0;
  double pa[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (NSUM > 0 && wave < NSUM) {
    const double* p = A.part + (size_t)(FIRST + wave) * T.npart;
    for (int i0 = lane; i0 < T.npart; i0 += 8 * kWave) {
#pragma unroll
      for (int u = 0; u < 8; u++) pa[u] += p[i0 + u * kWave];
    }
  }
  // --- the scalars
  float alpha = 0.f, beta = 0.f;
  if constexpr (NSUM > 0) {
    __shared__ double tot[4];
    if (wave < NSUM) {
      const double v = group_sum<kWave>(((pa[0] + pa[1]) + (pa[2] + pa[3])) + ((pa[4] + pa[5]) + (pa[6] + pa[7])));
      if (lane == 0) tot[wave] = v;
    }
    __syncthreads();
    if constexpr (MODE == kTileDir) {
      const double rr = tot[0], rpo = tot[1], rz = tot[2], zq = tot[3];
      const double pqp = sc->pq_prev;
      const double be = pqp == 0.0 ? 0.0 : -zq / pqp;
      const double rp = rz + be * rpo;
      const bool bad = !(rr == rr) || !(rpo == rpo) || !(rz == rz) || !(zq == zq) || !(pqp == pqp);
      const bool lost = !(rz > 0.0) || !(rp > 0.0) || pqp < 0.0;
      int stop = 0;
      if (bad) stop = 3;
      else if (lost) stop = 6;
      else if (rr <= sc->rtol2bb) stop = 1;
      else if (rr <= A.delta2 * sc->rr_anchor) stop = 5;
      else if (A.j >= A.maxit) stop = 2;
      if (b == 0 && threadIdx.x == 0) {
        sc->iters = A.j;
        sc->rr = rr;
        sc->rp = rp;
        if (stop) sc->stop = stop;
      }
      if (stop) return;
      beta = (float)be;
    } else {
      const double pq = tot[0];
      const double rp = sc->rp;
      alpha = (pq > 0.0 && rp > 0.0) ? (float)(rp / pq) : 0.f;
      if (b == 0 && threadIdx.x == 0) sc->pq_prev = pq;
    }
  }
  // --- the tile, the own rows' operands and the entry block into LDS
#pragma unroll
  for (int u = 0; u < TS; u++) {
    const int e = threadIdx.x + u * kTileThreads;
    if (e >= L) continue;
    float t;
    if constexpr (MODE == kTileDir) t = va[u] + beta * vb[u];
    else if constexpr (MODE == kTileSmooth) t = A.omega * vc[u] * (va[u] - alpha * vb[u]);
    else if constexpr (MODE == kTilePre) t = A.omega * vc[u] * va[u];
    else t = va[u];
    tile[e] = in[u] ? t : 0.f;
    const int o = own[u];
    if (o >= 0) {
      if constexpr (MODE == kTileSmooth) { e0s[o] = va[u] - alpha * vb[u]; e1s[o] = vd[u]; e2s[o] = ve[u]; }
      else if constexpr (MODE == kTilePre) { e0s[o] = va[u]; e1s[o] = vd[u]; }
      else if constexpr (MODE == kTilePost) { e0s[o] = vb[u]; e1s[o] = vc[u]; e2s[o] = vd[u]; }
    }
  }
  if constexpr (TEN) {
    if (threadIdx.x < ncf) cxs[threadIdx.x] = cf;   // ncf <= 1024 (nx + ty + tz <= 102)
  } else {
#pragma unroll
    for (int u = 0; u < NE; u++) {
      const int q = threadIdx.x + u * kTileThreads;
      if (q < T.estride) tes[q] = eb[u];
    }
  }
  __syncthreads();
  // --- this thread's row
  const int el = rw.w & 0xff;   // 0 past the tile's rows
  const bool act = el > 0;
  // (a lane past the tile's rows reads around a valid node in the uniform path)
  const int centre = kTileHalo + T.lx * (kTileHalo + T.ly * kTileHalo);
  const int grow = rw.x, li = act ? (rw.y & 0xffff) : centre, o = rw.y >> 16;
  float acc = 0.f;
  if constexpr (TEN) {
    // z pass: every (x, y) of the tile at the own z
    const int lxy = T.lx * T.ly;
    for (int q = threadIdx.x; q < n1; q += kTileThreads) {
      const int zo = q / lxy, xy = q - zo * lxy;
      const float* src = tile + xy + lxy * zo;   // tile z = zo + d, d = 0..4
      const float* cz = czs + 10 * zo;
      float a = 0.f, bb = 0.f;
#pragma unroll
      for (int d = 0; d < 5; d++) {
        const float t = src[d * lxy];
        bb += cz[d] * t;
        a += cz[5 + d] * t;
      }
      a1s[q] = a;
      b1s[q] = bb;
    }
    __syncthreads();
    // y pass: every x at the own y and z
    const int lxt = T.lx * T.ty;
    for (int q = threadIdx.x; q < n2; q += kTileThreads) {
      const int zo = q / lxt, r2 = q - zo * lxt, yo = r2 / T.lx, x = r2 - yo * T.lx;
      const int base = x + T.lx * (yo + T.ly * zo);   // products' y = yo + d, d = 0..4
      const float* cy = cys + 10 * yo;
      float c = 0.f, de = 0.f;
#pragma unroll
      for (int d = 0; d < 5; d++) {
        const float a = a1s[base + d * T.lx], bb = b1s[base + d * T.lx];
        c += cy[5 + d] * a;
        de += cy[d] * a + cy[5 + d] * bb;
      }
      c2s[q] = c;
      d2s[q] = de;
    }
    __syncthreads();
    // x pass at this thread's row: own index o = x + nx (yo + ty zo)
    if (act) {
      const int yz = o / T.nx, x = o - yz * T.nx;
      const int base = x + T.lx * yz;   // products' x = x + d, d = 0..4
      const float* cx = cxs + 10 * x;
#pragma unroll
      for (int d = 0; d < 5; d++) acc += cx[d] * c2s[base + d] + cx[5 + d] * d2s[base + d];
      if (grow == T.zp) acc += T.pen * tile[li];
    }
  } else {
  const int es0 = __builtin_amdgcn_readfirstlane(rw.z);
  const unsigned long long mixed = __ballot(act && rw.z != es0);
  if (mixed == 0) {
    // one class in the wavefront (rows past the tile's end have no entries):
    // its entries through scalar loads
    const int els = __builtin_amdgcn_readfirstlane(el);
    const int2* __restrict__ e = ent + es0;
    int k = 0;
    for (; k + 8 <= els; k += 8) {
      int2 t[8];
#pragma unroll
      for (int u = 0; u < 8; u++) t[u] = e[k + u];
#pragma unroll
      for (int u = 0; u < 8; u++) acc += __int_as_float(t[u].y) * tile[li + t[u].x];
    }
    for (; k < els; k++) {
      const int2 t = e[k];
      acc += __int_as_float(t.y) * tile[li + t.x];
    }
  } else {
    // rows of several classes: each lane reads its class's entries from LDS
    const int2* e = tes + (rw.w >> 8);
    for (int k = 0; k < el; k++) {
      const int2 t = e[k];
      acc += __int_as_float(t.y) * tile[li + t.x];
    }
  }
  }
  double s0 = 0.0, s1 = 0.0;
  if (act) {
    if constexpr (MODE == kTileDir) {
      const float pv = tile[li];
      A.pout[grow] = pv;
      A.q[grow] = acc;
      s0 = (double)pv * (double)acc;
    } else if constexpr (MODE == kTileSmooth) {
      const float rn = e0s[o], pv = e1s[o];
      A.rout[grow] = rn;
      A.x[grow] = e2s[o] + alpha * pv;
      A.r1[grow] = rn - acc;
      s0 = (double)rn * (double)rn;
      s1 = (double)rn * (double)pv;
    } else if constexpr (MODE == kTilePre) {
      const float rn = e0s[o];
      A.r1[grow] = rn - acc;
      s0 = (double)rn * (double)rn;
      s1 = (double)rn * (double)e1s[o];
    } else {
      const float rn = e0s[o];
      const float zv = tile[li] + A.omega * e1s[o] * (rn - acc);
      A.zout[grow] = zv;
      s0 = (double)rn * (double)zv;
      s1 = (double)zv * (double)e2s[o];
    }
  }
  double v[2] = {s0, s1};
  __shared__ double red[2 * (kTileThreads / kWave)];
  block_sum_all<2>(v, red);
  if (threadIdx.x == 0) {
    const int slot = MODE == kTileDir ? 0 : (MODE == kTilePost ? 3 : 1);
    A.part[slot * T.npart + b] = v[0];
    if (MODE != kTileDir) A.part[(slot + 1) * T.npart + b] = v[1];
  }
}

// A replaced residual (reliable update): r = float(rh), x_lo = 0, and the
// partials of r.r (slot 1), r.p (slot 2) and r.z (slot 3) for the kept
// direction p and the kept z; z.q (slot 4) is unchanged. z was formed from the
// FP32 recurrence residual, which the replacement moves by its rounding drift
// only, and flexible CG takes a preconditioner that varies that little. Block b
// sums its share of the nodes, so the partials keep the tile kernels' layout.
__global__ __launch_bounds__(kBlock) void k_tile_rescan(int n, int ntiles, int npart, const double* __restrict__ rh,
                                                        float* __restrict__ r, float* __restrict__ x,
                                                        const float* __restrict__ p, const float* __restrict__ z,
                                                        double* __restrict__ part) {
  const int b = blockIdx.x;
  const long long lo = (long long)n * b / ntiles, hi = (long long)n * (b + 1) / ntiles;
  double v[3] = {0.0, 0.0, 0.0};
  for (long long i = lo + threadIdx.x; i < hi; i += kBlock) {
    const float ri = (float)rh[i];
    r[i] = ri;
    x[i] = 0.0f;
    v[0] += (double)ri * (double)ri;
    v[1] += (double)ri * (double)p[i];
    v[2] += (double)ri * (double)z[i];
  }
  __shared__ double red[3 * (kBlock / kWave)];
  block_sum_all<3>(v, red);
  if (threadIdx.x == 0)
    for (int s = 0; s < 3; s++) part[(1 + s) * (size_t)npart + b] = v[s];
}

int launch_tile_rescan(fsx_ctx* ctx, const double* rh, float* r, const float* p) {
  const TilePlan& P = ctx->tile;
  hipLaunchKernelGGL(k_tile_rescan, dim3(P.ntiles), dim3(kBlock), 0, ctx->stream, ctx->NNp, P.ntiles, P.npart, rh, r,
                     ctx->fx, p, ctx->fz, ctx->tile_part);
  return launch_check(ctx, "k_tile_rescan");
}

template <int MODE, bool TEN>
static void launch_tile_ts(fsx_ctx* ctx, const TileGeom& T, const TileVecs& v, size_t lds) {
  const TilePlan& P = ctx->tile;
  const int L = T.lx * T.ly * T.lz;
  const dim3 grid(T.ntiles), block(kTileThreads);
  if (L <= 5 * kTileThreads)
    hipLaunchKernelGGL((k_tile<MODE, 5, TEN>), grid, block, lds, ctx->stream, T, P.rows, P.ent, P.tent, v);
  else if (L <= 9 * kTileThreads)
    hipLaunchKernelGGL((k_tile<MODE, 9, TEN>), grid, block, lds, ctx->stream, T, P.rows, P.ent, P.tent, v);
  else
    hipLaunchKernelGGL((k_tile<MODE, kTileSlotsMax, TEN>), grid, block, lds, ctx->stream, T, P.rows, P.ent, P.tent, v);
}

template <int MODE>
static void launch_tile_mode(fsx_ctx* ctx, const TileGeom& T, const TileVecs& v, size_t lds) {
  if (ctx->tile.tensor) launch_tile_ts<MODE, true>(ctx, T, v, lds);
  else launch_tile_ts<MODE, false>(ctx, T, v, lds);
}

int launch_tile(fsx_ctx* ctx, int mode, const TileVecs& v) {
  const TilePlan& P = ctx->tile;
  const TileGeom T{P.nx, P.ny, P.nz, P.ty, P.tz, P.lx, P.ly, P.lz, P.tiles_y, P.ntiles, P.npart, P.rstride, P.estride,
                   P.coef, P.zp, P.pen};
  const size_t L = (size_t)T.lx * T.ly * T.lz, nown = (size_t)T.nx * T.ty * T.tz;
  const size_t lds = P.tensor ? tensor_lds_floats(T.nx, T.ty, T.tz) * sizeof(float)
                              : (L + 3 * nown + ((L + 3 * nown) & 1)) * sizeof(float) + (size_t)T.estride * sizeof(int2);
  switch (mode) {
    case kTileDir: launch_tile_mode<kTileDir>(ctx, T, v, lds); break;
    case kTileSmooth: launch_tile_mode<kTileSmooth>(ctx, T, v, lds); break;
    case kTilePre: launch_tile_mode<kTilePre>(ctx, T, v, lds); break;
    default: launch_tile_mode<kTilePost>(ctx, T, v, lds); break;
  }
  return launch_check(ctx, "k_tile");
}

}  // namespace fsx
