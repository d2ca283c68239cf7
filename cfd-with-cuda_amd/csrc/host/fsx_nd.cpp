// fsx_nd.cpp -- nested-dissection order and front structure of Z for the
// multifrontal direct pressure solve.
//
// The MATLAB prototype orders the pressure matrix with symamd before its
// Cholesky factor (ML:1880-1881). Here the order is a nested dissection by
// recursive coordinate bisection: a node set larger than the leaf size is
// split at the median coordinate of its longest extent into A and B, and the
// smaller of {a in A coupled to B} and {b in B coupled to A} in Z's pattern
// is its separator, so no entry of Z couples the two remaining parts. The
// separator is a front whose children are the fronts of the parts; a set of
// at most the leaf size is a leaf front. Nodes are numbered in postorder
// (children's nodes first, then the front's own), so each front's pivots are
// a contiguous range and every front's boundary lies in its ancestors'
// separators. On the structured pressure grids the separators are the two
// node planes below a median plane (Z couples nodes two planes apart).
#include <algorithm>
#include <cmath>
#include <vector>

#include "../fsx_nd.h"
#include "fsx_host_internal.h"

namespace fsx {
namespace {

struct NdBuilder {
  int n = 0, leaf = 64;
  const int* rp = nullptr;
  const int* ci = nullptr;
  const double* xyz = nullptr;
  NdPlan* P = nullptr;
  std::vector<int> tag;   // 2 * stamp + side of the current split
  int stamp = 0;
  int next = 0;           // next new index

  int front(std::vector<int>& piv, int c0, int c1) {
    std::sort(piv.begin(), piv.end());
    NdFront f;
    f.p0 = next;
    f.npiv = (int)piv.size();
    f.child[0] = c0;
    f.child[1] = c1;
    for (int k = 0; k < f.npiv; k++) {
      P->perm[next + k] = piv[k];
      P->iperm[piv[k]] = next + k;
    }
    next += f.npiv;
    const int id = (int)P->fronts.size();
    for (int c : {c0, c1})
      if (c >= 0) P->fronts[c].parent = id;
    P->fronts.push_back(std::move(f));
    return id;
  }

  int build(std::vector<int>& S) {
    if ((int)S.size() <= leaf) return front(S, -1, -1);
    double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
    for (int v : S)
      for (int d = 0; d < 3; d++) {
        lo[d] = std::min(lo[d], xyz[3 * (size_t)v + d]);
        hi[d] = std::max(hi[d], xyz[3 * (size_t)v + d]);
      }
    int ax = 0;
    for (int d = 1; d < 3; d++)
      if (hi[d] - lo[d] > hi[ax] - lo[ax]) ax = d;
    if (!(hi[ax] > lo[ax])) return front(S, -1, -1);   // no extent left to split
    std::vector<double> c(S.size());
    for (size_t i = 0; i < S.size(); i++) c[i] = xyz[3 * (size_t)S[i] + ax];
    std::vector<double> cc = c;
    std::nth_element(cc.begin(), cc.begin() + cc.size() / 2, cc.end());
    const double med = cc[cc.size() / 2];
    std::vector<int> A, B;
    for (size_t i = 0; i < S.size(); i++) (c[i] < med ? A : B).push_back(S[i]);
    if (A.empty() || B.empty()) {
      A.clear(); B.clear();
      for (size_t i = 0; i < S.size(); i++) (c[i] <= med ? A : B).push_back(S[i]);
      if (A.empty() || B.empty()) return front(S, -1, -1);
    }
    std::vector<int>().swap(S);
    stamp++;
    const int ta = 2 * stamp, tb = 2 * stamp + 1;
    for (int v : A) tag[v] = ta;
    for (int v : B) tag[v] = tb;
    auto coupled = [&](int v, int other) {
      for (int e = rp[v]; e < rp[v + 1]; e++)
        if (tag[ci[e]] == other) return true;
      return false;
    };
    std::vector<int> sa, ra, sb, rb;
    for (int v : A) (coupled(v, tb) ? sa : ra).push_back(v);
    for (int v : B) (coupled(v, ta) ? sb : rb).push_back(v);
    std::vector<int> sep, p0, p1;
    if (sa.size() <= sb.size()) { sep.swap(sa); p0.swap(ra); p1.swap(B); }
    else { sep.swap(sb); p0.swap(A); p1.swap(rb); }
    std::vector<int>().swap(A); std::vector<int>().swap(B);
    std::vector<int>().swap(sa); std::vector<int>().swap(sb);
    std::vector<int>().swap(ra); std::vector<int>().swap(rb);
    const int c0 = p0.empty() ? -1 : build(p0);
    const int c1 = p1.empty() ? -1 : build(p1);
    return front(sep, c0, c1);
  }
};

}  // namespace

int nd_plan(int n, const int* rp, const int* ci, const double* xyz, int leaf, NdPlan* plan) {
  if (n <= 0 || !rp || !ci || !xyz || !plan || leaf < 1 || rp[0] != 0) return -1;
  for (int r = 0; r < n; r++) {
    if (rp[r + 1] < rp[r]) return -1;
    for (int e = rp[r]; e < rp[r + 1]; e++)
      if (ci[e] < 0 || ci[e] >= n) return -1;
  }
  for (size_t i = 0; i < 3 * (size_t)n; i++)
    if (!std::isfinite(xyz[i])) return -1;
  NdPlan& P = *plan;
  P = NdPlan{};
  P.perm.assign(n, -1);
  P.iperm.assign(n, -1);
  NdBuilder b;
  b.n = n; b.leaf = leaf; b.rp = rp; b.ci = ci; b.xyz = xyz; b.P = plan;
  b.tag.assign(n, 0);
  std::vector<int> all(n);
  for (int v = 0; v < n; v++) all[v] = v;
  b.build(all);
  // levels from the roots (created last)
  const int nf = (int)P.fronts.size();
  for (int t = nf - 1; t >= 0; t--) {
    NdFront& f = P.fronts[t];
    if (f.parent < 0) f.depth = 0;
    for (int c : f.child)
      if (c >= 0) P.fronts[c].depth = f.depth + 1;
    P.depth = std::max(P.depth, f.depth);
  }
  // boundaries, children first
  std::vector<int> seen(n, -1), list;
  for (int t = 0; t < nf; t++) {
    NdFront& f = P.fronts[t];
    const int hi = f.p0 + f.npiv;
    list.clear();
    for (int k = f.p0; k < hi; k++) {
      const int v = P.perm[k];
      for (int e = rp[v]; e < rp[v + 1]; e++) {
        const int j = P.iperm[ci[e]];
        if (j >= hi && seen[j] != t) { seen[j] = t; list.push_back(j); }
      }
    }
    for (int c : f.child)
      if (c >= 0)
        for (int j : P.fronts[c].bnd)
          if (j >= hi && seen[j] != t) { seen[j] = t; list.push_back(j); }
    std::sort(list.begin(), list.end());
    f.bnd = list;
    const double np = f.npiv, nb = (double)f.bnd.size(), m = np + nb;
    P.nnz_l += np * (np + 1) / 2 + np * nb;
    for (int j = 0; j < f.npiv; j++) P.flops += (m - j) * (m - j);
  }
  return 0;
}

}  // namespace fsx

int fsxh_nd_order(int n, const int* rp, const int* ci, const double* xyz, int leaf, int* perm, double* stats4) {
  fsx::NdPlan plan;
  if (fsx::nd_plan(n, rp, ci, xyz, leaf, &plan) != 0) return FSX_EINVAL;
  if (perm) std::copy(plan.perm.begin(), plan.perm.end(), perm);
  if (stats4) {
    stats4[0] = plan.nnz_l;
    stats4[1] = plan.flops;
    stats4[2] = (double)plan.fronts.size();
    stats4[3] = plan.depth;
  }
  return FSX_OK;
}
