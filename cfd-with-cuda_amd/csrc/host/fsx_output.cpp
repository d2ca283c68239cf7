// fsx_output.cpp -- field output and restart in the reference's Tecplot
// layout: createTecplot (BCH:4249-4482) and readRestartFile (BCH:4214-4242).
//
// File layout (ASCII, the one the reference writes and reads back):
//   TITLE = <title>
//   VARIABLES = x,  y,  z,  u, v, w, p
//   ZONE N=<NN>, E=<8 NE>, F=FEPOINT, ET=BRICK
//   NN lines "x y z u v w p" in %.11e (C++ scientific, precision 11), node
//     order of the global numbering; p at a non-corner node is interpolated
//     from the element's corners: mid-edge = mean of the edge's 2 corners,
//     mid-face = mean of the face's 4, centre = mean of all 8;
//   8 NE lines of 8 one-based node ids: each hex27 element split into its
//     8 octant sub-bricks (octant ox + 2 oy + 4 oz), vertices in hex8 order.
// A restart reads that file back: the first NCN node lines give u, v, w and
// p, the rest u, v, w.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "fsx_host_internal.h"

namespace {

// parametric position (-1, 0, +1) of each hex27 local node (BCH:2312-2343)
const int kPos[27][3] = {
  {-1,-1,-1},{1,-1,-1},{1,1,-1},{-1,1,-1},{-1,-1,1},{1,-1,1},{1,1,1},{-1,1,1},
  {0,-1,-1},{1,0,-1},{0,1,-1},{-1,0,-1},
  {-1,-1,0},{1,-1,0},{1,1,0},{-1,1,0},
  {0,-1,1},{1,0,1},{0,1,1},{-1,0,1},
  {0,0,-1},{0,-1,0},{1,0,0},{0,1,0},{-1,0,0},{0,0,1},{0,0,0}};

// local node at parametric position (a, b, c), each in {-1, 0, 1}
int local_at(int a, int b, int c) {
  for (int j = 0; j < 27; j++)
    if (kPos[j][0] == a && kPos[j][1] == b && kPos[j][2] == c) return j;
  return -1;
}

// sub[o][v]: local node of vertex v (hex8 order) of octant sub-brick o
void sub_bricks(int sub[8][8]) {
  const int hx[8] = {0, 1, 1, 0, 0, 1, 1, 0}, hy[8] = {0, 0, 1, 1, 0, 0, 1, 1}, hz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
  for (int o = 0; o < 8; o++)
    for (int v = 0; v < 8; v++)
      sub[o][v] = local_at((o & 1) + hx[v] - 1, ((o >> 1) & 1) + hy[v] - 1, ((o >> 2) & 1) + hz[v] - 1);
}

}  // namespace

// Writes the fields to `path`; Un is [u|v|w] (3 NN), Pn per corner node (NCN).
extern "C" int fsxh_write_tecplot(const fsxh_topo* t, const char* title, const double* Un, const double* Pn,
                                  const char* path) {
  if (!t || !Un || !Pn || !path) return FSX_EINVAL;
  const int NE = t->NE, NN = t->NN, NCN = t->NCN;
  std::vector<double> p(NN, 0.0);
  for (int i = 0; i < NCN; i++) p[i] = Pn[i];
  for (int e = 0; e < NE; e++) {
    const int* L = t->LtoGnode.data() + (size_t)e * 27;
    for (int ed = 0; ed < 12; ed++) p[L[8 + ed]] = 0.5 * (p[L[fsx::kEdge[ed][0]]] + p[L[fsx::kEdge[ed][1]]]);
    for (int f = 0; f < 6; f++)
      p[L[20 + f]] = 0.25 * (p[L[fsx::kFace[f][0]]] + p[L[fsx::kFace[f][1]]] + p[L[fsx::kFace[f][2]]] +
                             p[L[fsx::kFace[f][3]]]);
    p[L[26]] = 0.125 * (p[L[0]] + p[L[1]] + p[L[2]] + p[L[3]] + p[L[4]] + p[L[5]] + p[L[6]] + p[L[7]]);
  }
  FILE* fp = std::fopen(path, "w");
  if (!fp) return FSX_EINVAL;
  std::fprintf(fp, "TITLE = %s\n", title ? title : "");
  std::fprintf(fp, "VARIABLES = x,  y,  z,  u, v, w, p\n");
  std::fprintf(fp, "ZONE N=%d, E=%lld, F=FEPOINT, ET=BRICK\n", NN, 8LL * NE);
  for (int i = 0; i < NN; i++) {
    const double* x = t->coord.data() + 3 * (size_t)i;
    std::fprintf(fp, "%.11e %.11e %.11e %.11e %.11e %.11e %.11e\n", x[0], x[1], x[2], Un[i], Un[(size_t)NN + i],
                 Un[2 * (size_t)NN + i], p[i]);
  }
  int sub[8][8];
  sub_bricks(sub);
  for (int e = 0; e < NE; e++) {
    const int* L = t->LtoGnode.data() + (size_t)e * 27;
    for (int o = 0; o < 8; o++)
      std::fprintf(fp, "%d %d %d %d %d %d %d %d\n", L[sub[o][0]] + 1, L[sub[o][1]] + 1, L[sub[o][2]] + 1,
                   L[sub[o][3]] + 1, L[sub[o][4]] + 1, L[sub[o][5]] + 1, L[sub[o][6]] + 1, L[sub[o][7]] + 1);
  }
  const bool ok = std::ferror(fp) == 0;
  return (std::fclose(fp) == 0 && ok) ? FSX_OK : FSX_EINVAL;
}

// Reads a file in that layout into Un (3 NN) and Pn (NCN).
extern "C" int fsxh_read_restart(const fsxh_topo* t, const char* path, double* Un, double* Pn) {
  if (!t || !path || !Un || !Pn) return FSX_EINVAL;
  const int NN = t->NN, NCN = t->NCN;
  FILE* fp = std::fopen(path, "r");
  if (!fp) return FSX_EINVAL;
  std::vector<char> line(4096);
  int rc = FSX_OK;
  for (int k = 0; k < 3 && rc == FSX_OK; k++)
    if (!std::fgets(line.data(), (int)line.size(), fp)) rc = FSX_EINVAL;
  for (int i = 0; i < NN && rc == FSX_OK; i++) {
    double v[7];
    if (!std::fgets(line.data(), (int)line.size(), fp)) { rc = FSX_EINVAL; break; }
    char* s = line.data();
    for (int k = 0; k < 7 && rc == FSX_OK; k++) {
      char* end = nullptr;
      v[k] = std::strtod(s, &end);
      if (end == s) rc = FSX_EINVAL;
      s = end;
    }
    if (rc != FSX_OK) break;
    Un[i] = v[3]; Un[(size_t)NN + i] = v[4]; Un[2 * (size_t)NN + i] = v[5];
    if (i < NCN) Pn[i] = v[6];
  }
  std::fclose(fp);
  return rc;
}

// The driver's current Un, Pn to a Tecplot file (createTecplot, BCH:4249).
extern "C" int fsxd_write_tecplot(fsxd_solver* s, const char* title, const char* path) {
  if (!s || !s->ctx) return FSX_EINVAL;
  const size_t NN = s->topo->NN, NNp = s->topo->NNp;
  std::vector<double> un(3 * NN), pn(NNp);
  int rc = fsx_download(s->ctx, FSX_VEC_UN, un.data(), un.size());
  if (rc == FSX_OK) rc = fsx_download(s->ctx, FSX_VEC_PN, pn.data(), pn.size());
  if (rc != FSX_OK) { s->err = fsx_last_error(s->ctx); return rc; }
  rc = fsxh_write_tecplot(s->topo, title, un.data(), pn.data(), path);
  if (rc != FSX_OK) s->err = std::string("cannot write ") + path;
  return rc;
}

// Initial fields from a restart file instead of the initial BCs
// (readRestartFile, BCH:2795-2796): back to t_ini as fsxd_reset does, then
// Un and Pn replaced by the file's fields.
extern "C" int fsxd_read_restart(fsxd_solver* s, const char* path) {
  if (!s || !s->ctx) return FSX_EINVAL;
  const size_t NN = s->topo->NN, NNp = s->topo->NNp;
  std::vector<double> un(3 * NN, 0.0), pn(NNp, 0.0);
  int rc = fsxh_read_restart(s->topo, path, un.data(), pn.data());
  if (rc != FSX_OK) { s->err = std::string("cannot read restart file ") + path; return rc; }
  rc = fsxd_reset(s);
  if (rc != FSX_OK) return rc;
  rc = fsx_upload_vec(s->ctx, FSX_VEC_UN, un.data(), un.size());
  if (rc == FSX_OK) rc = fsx_upload_vec(s->ctx, FSX_VEC_PN, pn.data(), pn.size());
  if (rc != FSX_OK) s->err = fsx_last_error(s->ctx);
  return rc;
}
