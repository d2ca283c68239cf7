// fsx_gclass.cpp -- row classes of the gradient operator G and of G^T (the
// operators of step3, BCH:3912-3974, and of step2's right-hand side,
// BCH:3853-3881) for the step kernels at matrix precision 32.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "../fsx_gclass.h"
#include "fsx_host_internal.h"

namespace fsx {
namespace {

constexpr unsigned long long kFnvBasis = 1469598103934665603ull;
constexpr unsigned long long kFnvPrime = 1099511628211ull;

inline unsigned long long mix(unsigned long long h, unsigned long long v) { return (h ^ v) * kFnvPrime; }

inline unsigned bits(float f) {
  unsigned u;
  std::memcpy(&u, &f, sizeof u);
  return u;
}

// Interns lists of 32-bit words: equal lists get one id, new lists are
// appended to `store` (their words) with `ptr` their bounds.
struct ListDict {
  std::unordered_map<unsigned long long, std::vector<int>> by_hash;
  std::vector<int> ptr{0};
  std::vector<unsigned> store;

  int intern(const unsigned* w, int len) {
    unsigned long long h = mix(kFnvBasis, (unsigned long long)len);
    for (int k = 0; k < len; k++) h = mix(h, w[k]);
    auto& cand = by_hash[h];
    for (int c : cand) {
      if (ptr[c + 1] - ptr[c] != len) continue;
      if (std::equal(w, w + len, store.begin() + ptr[c])) return c;
    }
    const int id = (int)ptr.size() - 1;
    store.insert(store.end(), w, w + len);
    ptr.push_back((int)store.size());
    cand.push_back(id);
    return id;
  }
  int count() const { return (int)ptr.size() - 1; }
};

}  // namespace

int g_class_form(int nrows, int ncols, const int* rp, const int* ci, const double* g3, double prune_rel, int max_cls,
                 long long max_ent, int max_len, GClassForm* out) {
  if (nrows <= 0 || ncols <= 0 || !rp || !ci || !g3 || !out || rp[0] != 0 || !(prune_rel >= 0.0)) return -1;
  for (int r = 0; r < nrows; r++) {
    if (rp[r + 1] < rp[r]) return -1;
    for (int e = rp[r]; e < rp[r + 1]; e++)
      if (ci[e] < 0 || ci[e] >= ncols || (e > rp[r] && ci[e] <= ci[e - 1])) return -1;
  }
  GClassForm& F = *out;
  F = GClassForm{};
  const long long nnz = rp[nrows];
  // G: a row's key is its length, its column offsets and its FP32 values
  ListDict gd;
  std::vector<unsigned> key;
  F.row_base.assign(nrows, 0);
  F.row_cls.assign(nrows, 0);
  for (int r = 0; r < nrows; r++) {
    const int rs = rp[r], len = rp[r + 1] - rs;
    if (len > max_len) return 1;
    double m = 0.0;
    for (long long e = 3LL * rs; e < 3LL * (rs + len); e++) m = std::max(m, std::fabs(g3[e]));
    const double thr = m * prune_rel;
    key.resize(4 * (size_t)len);
    F.row_base[r] = len ? ci[rs] : 0;
    for (int k = 0; k < len; k++) {
      key[4 * k] = (unsigned)(ci[rs + k] - F.row_base[r]);
      for (int d = 0; d < 3; d++) {
        const double v = g3[3 * (size_t)(rs + k) + d];
        key[4 * k + 1 + d] = bits(std::fabs(v) < thr ? 0.0f : (float)v);
      }
    }
    F.row_cls[r] = gd.intern(key.data(), 4 * len);
    if (gd.count() > max_cls || (long long)gd.store.size() / 4 > max_ent) return 1;
  }
  const int ncls = gd.count();
  F.cls_ptr.resize(ncls + 1);
  for (int c = 0; c <= ncls; c++) F.cls_ptr[c] = gd.ptr[c] / 4;
  const size_t nent = gd.store.size() / 4;
  F.ent_off.resize(nent);
  F.ent_val.resize(3 * nent);
  for (size_t t = 0; t < nent; t++) {
    F.ent_off[t] = (int)gd.store[4 * t];
    for (int d = 0; d < 3; d++) std::memcpy(&F.ent_val[3 * t + d], &gd.store[4 * t + 1 + d], sizeof(float));
  }
  // G^T: every entry of G under its column, tagged with its class entry
  F.trow_ptr.assign(ncols + 1, 0);
  for (long long e = 0; e < nnz; e++) F.trow_ptr[ci[e] + 1]++;
  for (int p = 0; p < ncols; p++) F.trow_ptr[p + 1] += F.trow_ptr[p];
  std::vector<long long> ent(nnz);   // tag << 32 | row
  {
    std::vector<int> fill(F.trow_ptr.begin(), F.trow_ptr.end() - 1);
    for (int r = 0; r < nrows; r++)
      for (int e = rp[r]; e < rp[r + 1]; e++) {
        const long long tag = F.cls_ptr[F.row_cls[r]] + (e - rp[r]);
        ent[fill[ci[e]]++] = (tag << 32) | (unsigned)r;
      }
  }
  F.trow_cls.assign(ncols, 0);
  F.tcol.resize(nnz);
  for (int p = 0; p < ncols; p++) {
    const int ps = F.trow_ptr[p], len = F.trow_ptr[p + 1] - ps;
    if (len > max_len) return 1;
    std::sort(ent.begin() + ps, ent.begin() + ps + len);
    for (int k = 0; k < len; k++) F.tcol[ps + k] = (int)(ent[ps + k] & 0xffffffffLL);
  }
  // keyed by the tags (entries in tag order); then, where it makes no more
  // classes, by (column offset from the row's first column, FP32 values)
  // with the entries in column order, so that rows repeat their offsets
  // wherever the numbering does
  int ntag = -1;
  std::vector<long long> oe;
  for (int with_off = 0; with_off <= 1; with_off++) {
    ListDict td;
    bool fits = true;
    const int w = with_off ? 4 : 1;
    std::vector<int> cls(ncols), col;
    if (with_off) col.resize(nnz);
    for (int p = 0; p < ncols && fits; p++) {
      const int ps = F.trow_ptr[p], len = F.trow_ptr[p + 1] - ps;
      key.resize((size_t)w * len);
      if (with_off) {
        oe.resize(len);   // column << 32 | tag, ascending columns
        for (int k = 0; k < len; k++) oe[k] = ((long long)F.tcol[ps + k] << 32) | (ent[ps + k] >> 32);
        std::sort(oe.begin(), oe.end());
        const int c0 = (int)(oe[0] >> 32);
        for (int k = 0; k < len; k++) {
          col[ps + k] = (int)(oe[k] >> 32);
          const size_t tag = (size_t)(oe[k] & 0xffffffffLL);
          key[4 * (size_t)k] = (unsigned)(col[ps + k] - c0);
          for (int d = 0; d < 3; d++) key[4 * (size_t)k + 1 + d] = bits(F.ent_val[3 * tag + d]);
        }
      } else {
        for (int k = 0; k < len; k++) key[k] = (unsigned)(ent[ps + k] >> 32);
      }
      cls[p] = td.intern(key.data(), w * len);
      fits = td.count() <= max_cls && (long long)td.store.size() / w <= max_ent &&
             (!with_off || td.count() <= ntag);
    }
    if (!fits) {
      if (with_off) break;   // the tag-only classes stand
      return 1;
    }
    if (!with_off) ntag = td.count();
    else F.tcol.swap(col);
    F.trow_cls.swap(cls);
    const size_t ne = td.store.size() / w;
    F.toffsets = with_off != 0;
    F.tcls_ptr.resize(td.ptr.size());
    for (size_t c = 0; c < td.ptr.size(); c++) F.tcls_ptr[c] = td.ptr[c] / w;
    F.tent_val.resize(3 * ne);
    F.tent_off.assign(with_off ? ne : 0, 0);
    for (size_t t = 0; t < ne; t++) {
      if (with_off) {
        F.tent_off[t] = (int)td.store[4 * t];
        for (int d = 0; d < 3; d++) std::memcpy(&F.tent_val[3 * t + d], &td.store[4 * t + 1 + d], sizeof(float));
      } else {
        for (int d = 0; d < 3; d++) F.tent_val[3 * t + d] = F.ent_val[3 * (size_t)td.store[t] + d];
      }
    }
  }
  return 0;
}

}  // namespace fsx

int fsxh_g_classes(int nrows, int ncols, const int* rp, const int* ci, const double* g3, double prune_rel,
                   int* counts5, int* row_cls, int* trow_cls, float* g3_f32) {
  fsx::GClassForm F;
  const int rc = fsx::g_class_form(nrows, ncols, rp, ci, g3, prune_rel, 1 << 20, 1LL << 28, 1 << 20, &F);
  if (rc < 0) return FSX_EINVAL;
  if (counts5) {
    counts5[0] = rc == 0 ? (int)F.cls_ptr.size() - 1 : -1;
    counts5[1] = rc == 0 ? (int)F.ent_off.size() : -1;
    counts5[2] = rc == 0 ? (int)F.tcls_ptr.size() - 1 : -1;
    counts5[3] = rc == 0 ? (int)F.tcls_ptr.back() : -1;
    counts5[4] = rc == 0 ? (int)F.toffsets : -1;
  }
  if (rc != 0) return FSX_OK;
  if (row_cls) std::copy(F.row_cls.begin(), F.row_cls.end(), row_cls);
  if (trow_cls) std::copy(F.trow_cls.begin(), F.trow_cls.end(), trow_cls);
  if (g3_f32) {
    // G as the class form represents it: row r's entry k is class entry
    // cls_ptr[row_cls[r]] + k; G^T's rows are checked to hold the same values
    // under their columns, else the entry is written as NaN
    for (int r = 0; r < nrows; r++) {
      const int c0 = F.cls_ptr[F.row_cls[r]];
      for (int e = rp[r]; e < rp[r + 1]; e++)
        for (int d = 0; d < 3; d++) g3_f32[3 * (size_t)e + d] = F.ent_val[3 * (size_t)(c0 + e - rp[r]) + d];
    }
    for (int p = 0; p < ncols; p++) {
      const int c0 = F.tcls_ptr[F.trow_cls[p]];
      for (int k = F.trow_ptr[p]; k < F.trow_ptr[p + 1]; k++) {
        // with offsets in the classes, the column the class form reads
        const int r = F.toffsets ? F.tcol[F.trow_ptr[p]] + F.tent_off[c0 + k - F.trow_ptr[p]] : F.tcol[k];
        if (r < 0 || r >= nrows) return FSX_EINVAL;
        const int* lo = ci + rp[r];
        const int* hi = ci + rp[r + 1];
        const int* at = std::lower_bound(lo, hi, p);
        if (at == hi || *at != p) return FSX_EINVAL;   // a column of G^T that is no entry of G
        const size_t e = (size_t)(at - ci);
        for (int d = 0; d < 3; d++) {
          const float v = F.tent_val[3 * (size_t)(c0 + k - F.trow_ptr[p]) + d];
          if (fsx::bits(v) != fsx::bits(g3_f32[3 * e + d])) g3_f32[3 * e + d] = NAN;
        }
      }
    }
  }
  return FSX_OK;
}
