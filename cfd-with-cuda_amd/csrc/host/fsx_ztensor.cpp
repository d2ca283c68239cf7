// fsx_ztensor.cpp -- Z = sum_d G_d^T M^-1 G_d (BCH:3322-3513, CU:281-411)
// of a tensor-product box grid as Kronecker products of 1D stencils, for the
// learned solver's LDS-tiled Z products.
#include <algorithm>
#include <cmath>
#include <vector>

#include "../fsx_ztensor.h"
#include "fsx_host_internal.h"

namespace fsx {
namespace {

// Gauss-Legendre points and weights on [-1, 1] per axis of the NGP rule
// (the reference's 1-, 8- and 27-point tensor rules, BCH:2166-2220)
bool gauss_1d(int ngp, std::vector<double>& xi, std::vector<double>& w) {
  if (ngp == 1) { xi = {0.0}; w = {2.0}; return true; }
  if (ngp == 8) {
    const double a = 1.0 / std::sqrt(3.0);
    xi = {-a, a}; w = {1.0, 1.0};
    return true;
  }
  if (ngp == 27) {
    const double a = std::sqrt(0.6);
    xi = {-a, 0.0, a}; w = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
    return true;
  }
  return false;
}

// K and M of one axis with grid coordinates x[0..n-1]: five coefficients a
// point (offsets -2..2), K scaled by s
bool axis_tables(int n, const double* x, const std::vector<double>& xi, const std::vector<double>& w, double s,
                 std::vector<double>& K, std::vector<double>& M) {
  const int ne = n - 1, nv = 2 * ne + 1;
  std::vector<double> m(nv, 0.0), D(3 * 2 * (size_t)ne, 0.0), B(3 * 2 * (size_t)ne, 0.0);   // per element 3 x 2
  for (int e = 0; e < ne; e++) {
    const double h = x[e + 1] - x[e];
    for (size_t q = 0; q < xi.size(); q++) {
      const double t = xi[q];
      const double nn[3] = {0.5 * t * (t - 1.0), 1.0 - t * t, 0.5 * t * (t + 1.0)};
      const double dn[3] = {t - 0.5, -2.0 * t, t + 0.5};
      const double l[2] = {0.5 * (1.0 - t), 0.5 * (1.0 + t)};
      for (int a = 0; a < 3; a++) {
        m[2 * e + a] += w[q] * nn[a] * 0.5 * h;
        for (int p = 0; p < 2; p++) {
          D[(size_t)6 * e + 2 * a + p] += w[q] * dn[a] * l[p];
          B[(size_t)6 * e + 2 * a + p] += w[q] * nn[a] * l[p] * 0.5 * h;
        }
      }
    }
  }
  for (int a = 0; a < nv; a++)
    if (!(std::fabs(m[a]) > 0.0) || !std::isfinite(m[a])) return false;
  // (D^T m^-1 D)(p, q) = sum over velocity nodes a of D(a, p) D(a, q) / m(a);
  // node a's column entries: the elements that hold it
  K.assign(5 * (size_t)n, 0.0);
  M.assign(5 * (size_t)n, 0.0);
  for (int a = 0; a < nv; a++) {
    // (pressure node, D value, B value) of node a, from its one or two elements
    int pc[4];
    double dv[4], bv[4];
    int cnt = 0;
    for (int e = std::max(0, (a - 2 + 1) / 2); e <= std::min(ne - 1, a / 2); e++) {
      const int la = a - 2 * e;
      if (la < 0 || la > 2) continue;
      for (int p = 0; p < 2; p++) {
        pc[cnt] = e + p;
        dv[cnt] = D[(size_t)6 * e + 2 * la + p];
        bv[cnt] = B[(size_t)6 * e + 2 * la + p];
        cnt++;
      }
    }
    for (int i = 0; i < cnt; i++)
      for (int j = 0; j < cnt; j++) {
        const int off = pc[j] - pc[i];
        if (off < -2 || off > 2) return false;
        K[5 * (size_t)pc[i] + off + 2] += s * dv[i] * dv[j] / m[a];
        M[5 * (size_t)pc[i] + off + 2] += bv[i] * bv[j] / m[a];
      }
  }
  for (double v : K) if (!std::isfinite(v)) return false;
  for (double v : M) if (!std::isfinite(v)) return false;
  return true;
}

}  // namespace

int z_tensor_tables(int nx, int ny, int nz, const double* xc, int ngp, double rho, ZTensor* out) {
  if (!out || !xc || nx < 2 || ny < 2 || nz < 2 || !(rho > 0.0)) return 1;
  ZTensor& T = *out;
  T = ZTensor{};
  std::vector<double> xi, w;
  if (!gauss_1d(ngp, xi, w)) return 1;
  const int n[3] = {nx, ny, nz};
  const long long nxy = (long long)nx * ny, nn = nxy * nz;
  std::vector<double> ax[3];
  for (int d = 0; d < 3; d++) {
    ax[d].resize(n[d]);
    const long long stride = d == 0 ? 1 : (d == 1 ? nx : nxy);
    for (int i = 0; i < n[d]; i++) ax[d][i] = xc[3 * (i * stride) + d];
  }
  // separable coordinates: every node at (X[i], Y[j], Z[k])
  double ext = 0.0;
  for (int d = 0; d < 3; d++) ext = std::max(ext, std::fabs(ax[d][n[d] - 1] - ax[d][0]));
  const double tol = 1e-12 * (ext > 0.0 ? ext : 1.0);
  for (long long id = 0; id < nn; id++) {
    const long long idx[3] = {id % nx, (id / nx) % ny, id / nxy};
    for (int d = 0; d < 3; d++)
      if (!(std::fabs(xc[3 * id + d] - ax[d][idx[d]]) <= tol)) return 1;
  }
  const double s = 1.0 / (rho * rho);
  for (int d = 0; d < 3; d++) {
    T.n[d] = n[d];
    if (!axis_tables(n[d], ax[d].data(), xi, w, s, T.k[d], T.m[d])) return 1;
  }
  return 0;
}

int z_tensor_check(int n, const int* rp, const int* ci, const double* v, int zero_p, double tol, ZTensor* out) {
  ZTensor& T = *out;
  const int nx = T.n[0], ny = T.n[1], nz = T.n[2];
  if ((long long)nx * ny * nz != n || !rp || !ci || !v) return 1;
  T.zero_p = zero_p >= 0 && zero_p < n ? zero_p : -1;
  T.pen = 0.0;
  T.err = 0.0;
  std::vector<double> row(125);
  std::vector<char> seen(125);
  for (int r = 0; r < n; r++) {
    const int i = r % nx, j = (r / nx) % ny, k = r / (nx * ny);
    for (int c = 0; c < 125; c++) {
      const int dx = c % 5, dy = (c / 5) % 5, dz = c / 25;
      row[c] = T.k[0][5 * i + dx] * T.m[1][5 * j + dy] * T.m[2][5 * k + dz] +
               T.m[0][5 * i + dx] * T.k[1][5 * j + dy] * T.m[2][5 * k + dz] +
               T.m[0][5 * i + dx] * T.m[1][5 * j + dy] * T.k[2][5 * k + dz];
      seen[c] = 0;
    }
    double mx = 0.0;
    for (int e = rp[r]; e < rp[r + 1]; e++) mx = std::max(mx, std::fabs(v[e]));
    if (!(mx > 0.0)) return 1;
    double worst = 0.0;
    for (int e = rp[r]; e < rp[r + 1]; e++) {
      const int q = ci[e];
      const int di = q % nx - i, dj = (q / nx) % ny - j, dk = q / (nx * ny) - k;
      if (di < -2 || di > 2 || dj < -2 || dj > 2 || dk < -2 || dk > 2) return 1;
      const int c = (di + 2) + 5 * ((dj + 2) + 5 * (dk + 2));
      seen[c] = 1;
      double d = v[e] - row[c];
      if (r == T.zero_p && q == r) { T.pen = d; d = 0.0; }
      worst = std::max(worst, std::fabs(d));
    }
    for (int c = 0; c < 125; c++)
      if (!seen[c]) worst = std::max(worst, std::fabs(row[c]));
    T.err = std::max(T.err, worst / mx);
  }
  return T.err <= tol ? 0 : 1;
}

}  // namespace fsx

int fsxh_z_tensor(int nx, int ny, int nz, const double* xc, int ngp, double rho, int n, const int* rp, const int* ci,
                  const double* v, int zero_p, double* coef, double* info3) {
  if (!xc || !rp || !ci || !v || nx < 2 || ny < 2 || nz < 2 || (long long)nx * ny * nz != n) return FSX_EINVAL;
  fsx::ZTensor T;
  int rc = fsx::z_tensor_tables(nx, ny, nz, xc, ngp, rho, &T);
  const int sep = rc == 0;
  if (rc == 0) rc = fsx::z_tensor_check(n, rp, ci, v, zero_p, 1e-10, &T);
  if (info3) {
    info3[0] = sep ? T.err : -1.0;
    info3[1] = T.pen;
    info3[2] = rc == 0 ? 1.0 : 0.0;
  }
  if (coef && sep) {
    size_t at = 0;
    for (int d = 0; d < 3; d++)
      for (int i = 0; i < T.n[d]; i++) {
        for (int c = 0; c < 5; c++) coef[at++] = T.k[d][5 * (size_t)i + c];
        for (int c = 0; c < 5; c++) coef[at++] = T.m[d][5 * (size_t)i + c];
      }
  }
  return FSX_OK;
}
