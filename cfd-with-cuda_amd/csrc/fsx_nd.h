// fsx_nd.h -- nested-dissection order and front structure of the pressure
// operator Z for the multifrontal direct solve (host code, fsx_nd.cpp).
#pragma once

#include <vector>

namespace fsx {

// A front of the multifrontal factorization: its pivots are the new
// indices [p0, p0 + npiv); its rows are the pivots followed by the boundary
// (later-numbered indices coupled to the pivots in Z or to the children's
// boundaries), ascending.
struct NdFront {
  int p0 = 0, npiv = 0, depth = 0, parent = -1;
  int child[2] = {-1, -1};
  std::vector<int> bnd;
};

struct NdPlan {
  std::vector<int> perm;          // new index -> pressure id
  std::vector<int> iperm;         // pressure id -> new index
  std::vector<NdFront> fronts;    // postorder: children before parents, the roots last
  int depth = 0;                  // deepest level (roots are level 0)
  double nnz_l = 0.0;             // entries of L with dense fronts (pivot blocks' lower triangles)
  double flops = 0.0;             // factor flops, sum over L's columns of (column count)^2
};

// Recursive coordinate bisection of the n nodes (coordinates xyz, 3 a node)
// with vertex separators from Z's symmetric pattern (rp, ci; the diagonal
// may be present); sets of at most `leaf` nodes are leaf fronts. Returns 0,
// or -1 for inconsistent input (an index out of range, a row pointer that
// decreases).
int nd_plan(int n, const int* rp, const int* ci, const double* xyz, int leaf, NdPlan* plan);

}  // namespace fsx
