// fsx_ztensor.h -- the pressure operator Z of a tensor-product box grid as a
// sum of Kronecker products of 1D stencils (host code, fsx_ztensor.cpp).
#pragma once

#include <vector>

namespace fsx {

// On an nx x ny x nz grid of box elements whose corner coordinates separate
// (x depends on i only, y on j, z on k), hex27 velocity / hex8 pressure
// elements with tensor Gauss rules give G_x = -(1/rho) D_x (x) B_y (x) B_z
// (and alike), the row-sum lumped mass m_x (x) m_y (x) m_z, hence
//   Z = Kx (x) My (x) Mz + Mx (x) Ky (x) Mz + Mx (x) My (x) Kz,
//   K = D^T m^-1 D / rho^2,  M = B^T m^-1 B
// with D(a, p) = int n_a' l_p, B(a, p) = int n_a l_p, m(a) = int n_a over the
// 1D quadratic (n) and linear (l) bases of each axis. K and M couple grid
// points at most 2 apart: five coefficients per point, offsets -2..2.
struct ZTensor {
  int n[3] = {0, 0, 0};
  std::vector<double> k[3], m[3];   // n[d] x 5 per axis
  int zero_p = -1;                  // the pinned pressure node (-1: none)
  double pen = 0.0;                 // its diagonal beyond the tensor form (the pin's penalty)
  double err = 0.0;                 // largest mismatch against Z's rows, relative to the row's largest
};

// The axis tables from the corner coordinates (xc, 3 a node, node i + nx (j
// + ny k)), the Gauss points per axis (ngp 1, 8 or 27) and the density.
// Returns 0, or 1 when the coordinates do not separate or a table is not
// finite (no tensor form).
int z_tensor_tables(int nx, int ny, int nz, const double* xc, int ngp, double rho, ZTensor* out);

// Compares the tensor form with Z's rows (CSR rp/ci/v, the grid's node
// order; entries dropped as structural zeros absent) and records the pinned
// node's diagonal difference in out->pen. Returns 0 when every entry agrees
// within tol times its row's largest magnitude (tensor entries absent from
// the row included), else 1; out->err holds the largest mismatch.
int z_tensor_check(int n, const int* rp, const int* ci, const double* v, int zero_p, double tol, ZTensor* out);

}  // namespace fsx
