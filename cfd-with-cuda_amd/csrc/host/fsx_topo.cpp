// fsx_topo.cpp -- numbering, velocity BC nodes, monitor node, CSR patterns
// and element colouring for the fractional-step core (host, C++17 + OpenMP).
//
// The numbering reproduces the reference's setupNonCornerNodes (BCH:954-1320)
// without its coordinate search: a mid-edge (mid-face) node is created by the
// first (element, local edge) pair in element-major order whose edge (face) has
// not been seen, exactly the pair at which the reference finds no matching
// neighbour node; here edges and faces are identified by their sorted corner
// ids (sort-based, O(NE log NE)) instead of by coordinates within 1e-10.
#include "fsx_host_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>

#include <omp.h>

namespace fsx {
const int kEdge[12][2] = {{0,1},{1,2},{2,3},{3,0},{0,4},{1,5},{2,6},{3,7},{4,5},{5,6},{6,7},{7,4}};  // BCH:1002-1054
const int kFace[6][4] = {{0,1,2,3},{0,1,4,5},{1,2,5,6},{2,3,6,7},{0,3,4,7},{4,5,6,7}};                // BCH:1166-1204
const int kFaceMid[6][5] = {{8,9,10,11,20},{8,12,13,16,21},{9,13,14,17,22},
                            {10,14,15,18,23},{11,12,15,19,24},{16,17,18,19,25}};                        // BCH:1480-1527
}  // namespace fsx

using namespace fsx;

namespace {

struct Key4 { int a, b, c, d; int idx; };

// Assign ids to the entities (edges or faces) listed per element slot: the
// first slot (in slot order) of each distinct key creates a new node.
template <class K, class Less, class Eq>
int assign_new_nodes(std::vector<K>& keys, int first_id, std::vector<int>& id_of_slot, Less less, Eq eq) {
  const size_t n = keys.size();
  std::sort(keys.begin(), keys.end(), less);           // by key, then by slot
  std::vector<int> creator(n);                         // creator slot of each sorted entry
  std::vector<char> is_creator(n, 0);
  size_t g = 0;
  for (size_t i = 0; i < n; i++) {
    if (i == 0 || !eq(keys[i], keys[i-1])) g = i;
    creator[i] = keys[g].idx;
    if (g == i) is_creator[keys[i].idx] = 1;
  }
  std::vector<int> new_id(n, -1);
  int next = first_id;
  for (size_t s = 0; s < n; s++) if (is_creator[s]) new_id[s] = next++;
  id_of_slot.assign(n, -1);
  for (size_t i = 0; i < n; i++) id_of_slot[keys[i].idx] = new_id[creator[i]];
  return next;
}

void incidence(int nnodes, int NE, const std::vector<int>& L, int per_elem, std::vector<int>& ptr, std::vector<int>& el) {
  ptr.assign(nnodes + 1, 0);
  for (int e = 0; e < NE; e++) for (int i = 0; i < per_elem; i++) ptr[L[(size_t)e*27 + i] + 1]++;
  for (int n = 0; n < nnodes; n++) ptr[n + 1] += ptr[n];
  el.resize(ptr[nnodes]);
  std::vector<int> next(ptr.begin(), ptr.end() - 1);
  for (int e = 0; e < NE; e++) for (int i = 0; i < per_elem; i++) el[next[L[(size_t)e*27 + i]]++] = e;
}

// Row r: sorted union of the first `ncol` local nodes of the elements at r
// (setupSparseM BCH:1700-1770 with ncol 27, setupSparseG BCH:1948-2010 with 8).
void pattern(const fsxh_topo& t, const std::vector<int>& vptr, const std::vector<int>& vel, int ncol,
             std::vector<int>& rowptr, std::vector<int>& col) {
  const int NN = t.NN;
  rowptr.assign(NN + 1, 0);
  std::vector<int> cnt(NN);
  #pragma omp parallel
  {
    std::vector<int> buf;
    #pragma omp for schedule(static)
    for (int r = 0; r < NN; r++) {
      buf.clear();
      for (int k = vptr[r]; k < vptr[r+1]; k++) {
        const int* L = &t.LtoGnode[(size_t)vel[k]*27];
        buf.insert(buf.end(), L, L + ncol);
      }
      std::sort(buf.begin(), buf.end());
      cnt[r] = (int)(std::unique(buf.begin(), buf.end()) - buf.begin());
    }
  }
  for (int r = 0; r < NN; r++) rowptr[r+1] = rowptr[r] + cnt[r];
  col.resize((size_t)rowptr[NN]);
  #pragma omp parallel
  {
    std::vector<int> buf;
    #pragma omp for schedule(static)
    for (int r = 0; r < NN; r++) {
      buf.clear();
      for (int k = vptr[r]; k < vptr[r+1]; k++) {
        const int* L = &t.LtoGnode[(size_t)vel[k]*27];
        buf.insert(buf.end(), L, L + ncol);
      }
      std::sort(buf.begin(), buf.end());
      std::unique(buf.begin(), buf.end());
      std::copy(buf.begin(), buf.begin() + cnt[r], col.begin() + rowptr[r]);
    }
  }
}

}  // namespace

extern "C" int fsxh_topology(const fsxh_mesh* m, fsxh_topo** out, char* err, int errlen) {
  auto fail = [&](const std::string& msg) {
    if (err && errlen > 0) std::snprintf(err, errlen, "%s", msg.c_str());
    return (int)FSX_EINVAL;
  };
  auto t = std::make_unique<fsxh_topo>();
  const int NE = m->NE, NCN = m->NCN;
  t->NE = NE; t->NCN = NCN; t->NNp = NCN; t->zeroP = m->zeroP;
  t->LtoGnode.assign((size_t)NE * 27, -1);
  for (int e = 0; e < NE; e++)
    for (int i = 0; i < 8; i++) {
      int n = m->conn[(size_t)e*8 + i];
      if (n < 0 || n >= NCN) return fail("corner index out of range");
      t->LtoGnode[(size_t)e*27 + i] = n;
    }
  // ---- mid-edge nodes (BCH:977-1134)
  std::vector<Key4> ek((size_t)NE * 12);
  for (int e = 0; e < NE; e++)
    for (int ed = 0; ed < 12; ed++) {
      int a = t->LtoGnode[(size_t)e*27 + kEdge[ed][0]], b = t->LtoGnode[(size_t)e*27 + kEdge[ed][1]];
      ek[(size_t)e*12 + ed] = Key4{std::min(a, b), std::max(a, b), 0, 0, e*12 + ed};
    }
  auto less = [](const Key4& x, const Key4& y) {
    if (x.a != y.a) return x.a < y.a;
    if (x.b != y.b) return x.b < y.b;
    if (x.c != y.c) return x.c < y.c;
    if (x.d != y.d) return x.d < y.d;
    return x.idx < y.idx;
  };
  auto eq = [](const Key4& x, const Key4& y) { return x.a == y.a && x.b == y.b && x.c == y.c && x.d == y.d; };
  std::vector<int> slot_id;
  int count = assign_new_nodes(ek, NCN, slot_id, less, eq);
  const int firstFace = count;
  for (int e = 0; e < NE; e++) for (int ed = 0; ed < 12; ed++) t->LtoGnode[(size_t)e*27 + 8 + ed] = slot_id[(size_t)e*12 + ed];
  ek.clear(); ek.shrink_to_fit();
  // ---- mid-face nodes (BCH:1140-1228)
  std::vector<Key4> fk((size_t)NE * 6);
  for (int e = 0; e < NE; e++)
    for (int f = 0; f < 6; f++) {
      int v[4];
      for (int k = 0; k < 4; k++) v[k] = t->LtoGnode[(size_t)e*27 + kFace[f][k]];
      std::sort(v, v + 4);
      fk[(size_t)e*6 + f] = Key4{v[0], v[1], v[2], v[3], e*6 + f};
    }
  count = assign_new_nodes(fk, count, slot_id, less, eq);
  for (int e = 0; e < NE; e++) for (int f = 0; f < 6; f++) t->LtoGnode[(size_t)e*27 + 20 + f] = slot_id[(size_t)e*6 + f];
  fk.clear(); fk.shrink_to_fit();
  // ---- centre nodes (BCH:1235-1259)
  const int firstCentre = count;
  for (int e = 0; e < NE; e++) t->LtoGnode[(size_t)e*27 + 26] = count++;
  t->NN = count;
  const int NN = t->NN;
  // ---- coordinates of every node, written by its creating element with the
  // reference's formulas (BCH:1058-1060, 1206-1208, 1246-1248)
  t->coord.assign((size_t)NN * 3, 0.0);
  std::copy(m->coord.begin(), m->coord.end(), t->coord.begin());
  std::vector<char> done((size_t)NN, 0);
  for (int e = 0; e < NE; e++) {
    const int* L = &t->LtoGnode[(size_t)e*27];
    auto X = [&](int n, int d) { return t->coord[3*(size_t)n + d]; };
    for (int ed = 0; ed < 12; ed++) {
      int n = L[8 + ed];
      if (done[n]) continue;
      done[n] = 1;
      int n1 = L[kEdge[ed][0]], n2 = L[kEdge[ed][1]];
      for (int d = 0; d < 3; d++) t->coord[3*(size_t)n + d] = 0.5 * (X(n1, d) + X(n2, d));
    }
  }
  for (int e = 0; e < NE; e++) {
    const int* L = &t->LtoGnode[(size_t)e*27];
    auto X = [&](int n, int d) { return t->coord[3*(size_t)n + d]; };
    for (int f = 0; f < 6; f++) {
      int n = L[20 + f];
      if (done[n]) continue;
      done[n] = 1;
      int a = L[kFace[f][0]], b = L[kFace[f][1]], c = L[kFace[f][2]], d4 = L[kFace[f][3]];
      for (int d = 0; d < 3; d++) t->coord[3*(size_t)n + d] = 0.25 * (X(a, d) + X(b, d) + X(c, d) + X(d4, d));
    }
    int n = L[26];
    for (int d = 0; d < 3; d++) {
      double s = 0.0;
      for (int k = 0; k < 8; k++) s += X(L[k], d);
      t->coord[3*(size_t)n + d] = 0.125 * s;
    }
  }
  (void)firstFace; (void)firstCentre;
  // ---- velocity BC nodes (BCH:1410-1580): last face in file order wins
  std::vector<int> info(NN, -1);
  const size_t nf = m->faces.size() / 3;
  for (size_t i = 0; i < nf; i++) {
    int e = m->faces[3*i], f = m->faces[3*i+1], w = m->faces[3*i+2];
    if (e < 0 || e >= NE || f < 0 || f > 5 || w < 0 || w >= m->nBC) return fail("BC face out of range");
    for (int k = 0; k < 4; k++) info[t->LtoGnode[(size_t)e*27 + kFace[f][k]]] = w;
    for (int k = 0; k < 5; k++) info[t->LtoGnode[(size_t)e*27 + kFaceMid[f][k]]] = w;
  }
  for (int i = 0; i < NN; i++)
    if (info[i] != -1) {
      t->BCnodes.push_back(i);
      t->BCwhich.push_back(info[i]);
      double u = m->BCstr[3*info[i]], v = m->BCstr[3*info[i]+1], w = m->BCstr[3*info[i]+2];
      if (m->has_profile && info[i] == m->profile_bc) {
        double s = t->coord[3*(size_t)i+1] - m->profile_y0, tt = t->coord[3*(size_t)i+2] - m->profile_z0;
        u = m->profile_peak * 16.0 * s * (1.0 - s) * tt * (1.0 - tt);
        v = 0.0; w = 0.0;
      }
      t->BCvalues.push_back(u); t->BCvalues.push_back(v); t->BCvalues.push_back(w);
    }
  // ---- monitor node (BCH:1644-1668)
  double dist = 1e6;
  for (int i = 0; i < NCN; i++) {
    double dx = t->coord[3*(size_t)i] - m->mon[0], dy = t->coord[3*(size_t)i+1] - m->mon[1], dz = t->coord[3*(size_t)i+2] - m->mon[2];
    double r = std::sqrt(dx*dx + dy*dy + dz*dz);
    if (r < dist) { dist = r; t->monPoint = i; }
  }
  // ---- CSR patterns
  std::vector<int> vptr, vel;
  incidence(NN, NE, t->LtoGnode, 27, vptr, vel);
  pattern(*t, vptr, vel, 27, t->Mrow, t->Mcol);
  pattern(*t, vptr, vel, 8, t->Grow, t->Gcol);
  if ((long long)t->Mrow[NN] > 2147483647LL) return fail("nnz(M) exceeds int32");
  // Z pattern: p ~ q iff some velocity row of G holds both (structural G^T G)
  {
    std::vector<int> pptr, pel;
    incidence(NCN, NE, t->LtoGnode, 8, pptr, pel);
    const int NNp = NCN;
    std::vector<std::vector<int>> rows(NNp);
    #pragma omp parallel
    {
      std::vector<int> markv(NN, -1), markp(NNp, -1);
      #pragma omp for schedule(dynamic, 256)
      for (int p = 0; p < NNp; p++) {
        std::vector<int>& r = rows[p];
        for (int k = pptr[p]; k < pptr[p+1]; k++) {
          const int* L = &t->LtoGnode[(size_t)pel[k]*27];
          for (int j = 0; j < 27; j++) {
            int i = L[j];
            if (markv[i] == p) continue;
            markv[i] = p;
            for (int g = t->Grow[i]; g < t->Grow[i+1]; g++) {
              int q = t->Gcol[g];
              if (markp[q] != p) { markp[q] = p; r.push_back(q); }
            }
          }
        }
        std::sort(r.begin(), r.end());
      }
    }
    t->Zrow.assign(NNp + 1, 0);
    for (int p = 0; p < NNp; p++) t->Zrow[p+1] = t->Zrow[p] + (int)rows[p].size();
    t->Zcol.resize((size_t)t->Zrow[NNp]);
    for (int p = 0; p < NNp; p++) std::copy(rows[p].begin(), rows[p].end(), t->Zcol.begin() + t->Zrow[p]);
    // ---- race-free colouring: greedy smallest free colour among elements
    // sharing a node (sharing any node implies sharing a corner)
    std::vector<int> color(NE, -1);
    int ncolors = 0;
    std::vector<int> seen(64, -1);
    for (int e = 0; e < NE; e++) {
      for (int i = 0; i < 8; i++) {
        int p = t->LtoGnode[(size_t)e*27 + i];
        for (int k = pptr[p]; k < pptr[p+1]; k++) {
          int c = color[pel[k]];
          if (c >= 0) seen[c] = e;
        }
      }
      int c = 0;
      while (c < 64 && seen[c] == e) c++;
      if (c >= 64) return fail("colouring needs more than 64 colours");
      color[e] = c;
      ncolors = std::max(ncolors, c + 1);
    }
    t->color_ptr.assign(ncolors + 1, 0);
    for (int e = 0; e < NE; e++) t->color_ptr[color[e] + 1]++;
    for (int c = 0; c < ncolors; c++) t->color_ptr[c + 1] += t->color_ptr[c];
    t->color_order.resize(NE);
    std::vector<int> next(t->color_ptr.begin(), t->color_ptr.end() - 1);
    for (int e = 0; e < NE; e++) t->color_order[next[color[e]]++] = e;
  }
  *out = t.release();
  return FSX_OK;
}

extern "C" void fsxh_topo_free(fsxh_topo* t) { delete t; }

extern "C" int fsxh_topo_info(const fsxh_topo* t, long long* out) {
  long long v[11] = {t->NE, t->NCN, t->NN, t->NNp, (long long)t->Mcol.size(), (long long)t->Gcol.size(),
                     (long long)t->Zcol.size(), (long long)t->BCnodes.size(), (long long)t->color_ptr.size() - 1,
                     t->monPoint, t->zeroP};
  std::memcpy(out, v, sizeof v);
  return FSX_OK;
}

static const std::vector<int>* topo_ivec(const fsxh_topo* t, const std::string& n) {
  if (n == "LtoGnode") return &t->LtoGnode;
  if (n == "Mrow") return &t->Mrow;
  if (n == "Mcol") return &t->Mcol;
  if (n == "Grow") return &t->Grow;
  if (n == "Gcol") return &t->Gcol;
  if (n == "Zrow") return &t->Zrow;
  if (n == "Zcol") return &t->Zcol;
  if (n == "BCnodes") return &t->BCnodes;
  if (n == "BCwhich") return &t->BCwhich;
  if (n == "color_ptr") return &t->color_ptr;
  if (n == "color_order") return &t->color_order;
  return nullptr;
}

extern "C" long long fsxh_topo_len(const fsxh_topo* t, const char* name) {
  std::string n(name);
  if (n == "coord") return (long long)t->coord.size();
  if (n == "BCvalues") return (long long)t->BCvalues.size();
  auto* v = topo_ivec(t, n);
  return v ? (long long)v->size() : -1;
}

extern "C" int fsxh_topo_get(const fsxh_topo* t, const char* name, void* dst) {
  std::string n(name);
  if (n == "coord") { std::memcpy(dst, t->coord.data(), t->coord.size() * sizeof(double)); return FSX_OK; }
  if (n == "BCvalues") { std::memcpy(dst, t->BCvalues.data(), t->BCvalues.size() * sizeof(double)); return FSX_OK; }
  auto* v = topo_ivec(t, n);
  if (!v) return FSX_EINVAL;
  std::memcpy(dst, v->data(), v->size() * sizeof(int));
  return FSX_OK;
}
