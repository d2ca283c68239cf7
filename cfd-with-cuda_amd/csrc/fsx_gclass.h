// fsx_gclass.h -- the gradient operator G (velocity rows, pressure columns,
// three components an entry) and its transpose as row classes, for the step
// kernels at matrix precision 32 (host code, fsx_gclass.cpp).
#pragma once

#include <vector>

namespace fsx {

// Rows of G that take the same list of (column offset from the row's first
// column, FP32 value triple) share one class: on a mesh of congruent
// elements every interior row of one node kind is the same list, so the
// step kernels read a class entry (cached) and the row's first column in
// place of each row's columns and values.
//
// The rows of G^T are classed through G's: entry (r, k) of G (row r, its
// k-th entry) carries the tag cls_ptr[row_cls[r]] + k, the global index of
// that entry in its class; a row of G^T lists its entries in ascending tag
// order (ties by row), and rows with the same tag sequence share one class,
// whose values are those of the tags' class entries. Where the rows' column
// offsets from their first (in that order) repeat too, they join the key
// and the class entries hold them (`toffsets`, a velocity numbering in grid
// order); otherwise the columns stay per row, each row in its class's order.
struct GClassForm {
  // G
  std::vector<int> row_base;    // first column of each row
  std::vector<int> row_cls;     // class of each row
  std::vector<int> cls_ptr;     // class c's entries [cls_ptr[c], cls_ptr[c+1])
  std::vector<int> ent_off;     // column offset from the row's first column
  std::vector<float> ent_val;   // (G1, G2, G3) an entry
  // G^T (rows are the columns of G)
  std::vector<int> trow_ptr;    // G^T's row pointers (entries of G per column)
  std::vector<int> trow_cls;    // class of each row of G^T
  std::vector<int> tcol;        // G^T's columns, each row in its class's order
  std::vector<int> tcls_ptr;    // class c's entries [tcls_ptr[c], tcls_ptr[c+1])
  std::vector<float> tent_val;  // (G1, G2, G3) an entry
  bool toffsets = false;        // the classes carry the column offsets:
  std::vector<int> tent_off;    //   column offset from the row's first column, per class entry
};

// Builds the classes of G (nrows rows over ncols columns, CSR rp/ci with
// ascending columns, g3 = (G1, G2, G3) an entry in FP64). A component below
// prune_rel times the largest magnitude of its row is taken as 0 (rounding
// residue of cancelling element contributions); the others are rounded to
// FP32. Returns 0 with the classes in *out, 1 when the rows do not compress
// (more than max_cls classes of either operator, more than max_ent class
// entries, or a row longer than max_len), -1 for inconsistent input.
int g_class_form(int nrows, int ncols, const int* rp, const int* ci, const double* g3, double prune_rel, int max_cls,
                 long long max_ent, int max_len, GClassForm* out);

}  // namespace fsx
