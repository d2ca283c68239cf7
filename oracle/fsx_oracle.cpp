// fsx_oracle.cpp -- CPU ORACLE. TEST INFRASTRUCTURE ONLY.
//
// This file is the checker for the HIP product path, never part of it: only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load the
// library built from it (oracle/_build/libfsx_oracle.so).
//
// It restates, in plain C++17 + OpenMP, the CPU path of the reference solver
// fractionalStep/Cpp/blascoCodinaHuerta.cpp (cited below as BCH:<line>):
//   * .inp reader                       BCH:528-725
//   * element/node topology             BCH:732-846
//   * greedy mesh colouring (with the reference's i=0 reset quirk) BCH:853-947
//   * mid-edge/mid-face/centre node creation by coordinate matching BCH:954-1320
//   * LtoGvel layout [u|v|w]            BCH:1327-1403
//   * face BCs -> node BCs, last writer wins BCH:1410-1580
//   * CSR patterns of M/K/A and G and the local->CSR maps BCH:1675-2159
//   * Gauss rule (NGP 1/8 as in BCH:2166-2208, plus the 27-point tensor rule
//     that the reference leaves as a TODO), hex27/hex8 shape functions
//     BCH:2215-2488, Jacobians BCH:2495-2711
//   * step0 static assembly, lumped mass, inverses BCH:3127-3315
//   * Z = sum_d G_d^T diag(MdOrigInv) G_d by a column SpGEMM in the
//     accumulation order of CSparse's cs_multiply, penalty pin BCH:3322-3479,
//     BCH:4154-4186
//   * coloured convection residual BCH:3520-3685
//   * step1/step2/step3 BCH:3692-3974, K*Acc_prev BCH:2981-3035
//   * Jacobi-PCG BCH:3981-4055 (textbook recurrence, x0 = 0, stop on
//     ||r||_2 <= rtol ||b||_2; MKL's RCI internals are not public source)
//   * time loop, inner convergence and steady-state checks BCH:2815-3120
//
// Deliberate, documented differences from the reference (none changes a
// result):
//   * gDSv is recomputed per element instead of stored (same arithmetic);
//   * the local->CSR map uses binary search in the sorted row instead of a
//     linear search (same index, the row is sorted and duplicate free);
//   * the convection scatter is done sequentially in colour order after a
//     parallel per-element compute, so the sum order is the reference's
//     colour order but no race is possible even where the colouring is
//     invalid (the i=0 reset quirk);
//   * SpMVs use the full symmetric Z instead of MKL's upper-triangle symv.
//
// Options with no reference counterpart, off by default, that model the
// device's other forms for the tests: consistent-mass updates (orc_set_mass),
// the warm-started PCG (orc_set_pcg_x0, the reference GPU path's x0),
// FP32-rounded operands (orc_set_f32_operands) and the direct Cholesky
// pressure solve of the MATLAB prototype (orc_set_direct, ML:1876-1885,
// 2021-2030).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cstdint>
#include <string>
#include <vector>
#include <algorithm>
#include <fstream>
#include <sstream>
#include <chrono>
#include <omp.h>

namespace {

struct Case {
  // ---- input (BCH:141-180) ----
  int NE = 0, NCN = 0, NN = 0, NNp = 0, NGP = 8;
  const int NENv = 27, NENp = 8, NEC = 8, NEE = 12, NEF = 6;
  double dt = 0.01, t_ini = 0.0, t_final = 1.0, tolerance = 1e-3, converge = 1e-6;
  double density = 1.0, viscosity = 0.01;
  int maxIter = 3;
  std::vector<double> coord;      // NN*3 after setup (NCN*3 before)
  std::vector<int> LtoGnode;      // NE*27, -1 before creation
  int nBC = 0;
  std::vector<double> BCstr;      // nBC*3
  std::vector<int> velFaces;      // nVelFaces*3 (elem, face, bc), 0-based
  int zeroP = -1;
  double mon[3] = {0.5, 0.5, 0.5};
  int monPoint = 0;
  // solver options (reference: rtol 1e-6 equivalent, maxit 1000)
  double pcg_rtol = 1e-8;
  int pcg_maxit = 1000;
  bool pcg_warm = false;          // x0 = previous Pdot (the GPU path, CU:425-442) instead of 0
  // FP32-rounded operands (lumped mass only): K's operands UnpHalf_prev and
  // Acc_prev before the K products of step1 and K*Acc, and step2's operand
  // UnpHalf / dt^2 - MdOrigInv KAcc before the G^T products: a model of the
  // device's FP32 operand shadows for tests, no reference counterpart
  bool f32_operands = false;
  // pressure solve by Cholesky of Z instead of PCG (the MATLAB prototype's
  // step, ML:1876-1885, 2021-2030); the factor is kept until Z changes
  bool pcg_direct = false;
  int chol_w = -1;                // half-bandwidth of the kept factor, -1: none
  std::vector<double> chol_L;     // lower band, row i at i*(w+1), column j at j - i + w
  // velocity mass: 0 lumped (the reference, BCH:3262-3295), 1 consistent M
  // solved by Jacobi-PCG for the three components (no reference counterpart)
  int mass_mode = 0;
  double mass_rtol = 1e-10;
  int mass_maxit = 1000;
  std::vector<double> Mbc;        // M with BC rows and columns replaced by the identity
  int last_mass_iters = 0;
  double last_mass_relres = 0.0;
  long total_mass_iters = 0;
  // ---- topology ----
  std::vector<int> elemsOfPres, nElemOfPres;   // NNp*10
  std::vector<int> neigh, nNeigh;              // NE*26
  std::vector<int> meshColors, nMeshColors, elementsOfColor;
  int nActiveColors = 0;
  std::vector<int> BCnodes, BCwhich;           // sorted by node
  std::vector<int> elemsOfVel, nElemOfVel;     // NN*10
  // ---- sparse ----
  std::vector<int> Mrow, Mcol;    // rowptr NN+1, col nnzM
  std::vector<int> Grow, Gcol;    // rowptr NN+1, col nnzG
  std::vector<int> mapM, mapG;    // NE*27*27, NE*27*8
  std::vector<double> Kval, Mval, G1, G2, G3;
  std::vector<int> Zrow, Zcol;
  std::vector<double> Zval;
  std::vector<int> Trow, Tcol, Tpos;  // G^T as CSR: rows = pressure nodes, ascending velocity rows
  // ---- reference element ----
  std::vector<double> gqp, gqw;   // NGP*3, NGP
  std::vector<double> Sv, Sp;     // NGP*27, NGP*8
  std::vector<double> dSv, dSp;   // [3][27][NGP], [3][8][NGP]
  // ---- lumped mass ----
  std::vector<double> Md, MdOrig, MdInv, MdOrigInv;  // 3NN
  // ---- state (BCH:2718-2808) ----
  std::vector<double> Un, Unp1, Unp1_prev, UnpHalf, UnpHalf_prev, Acc, Acc_prev, KAcc;
  std::vector<double> Pn, Pnp1, Pnp1_prev, Pdot, R1, R2, R3;
  int timeN = 0;
  double timeT = 0.0;
  // ---- statistics ----
  int last_pcg_iters = 0;
  double last_pcg_relres = 0.0;
  long total_pcg_iters = 0;
  int n_solves = 0;
  int last_inner_iters = 0;
  double maxAcc = 0.0;
  double t_step_total = 0.0, t_pcg_total = 0.0;
  std::string err;
};

// ---------------------------------------------------------------- .inp reader
// BCH:528-725. The reader skips to ':' for each header value, so the header
// line order is the contract; lines are otherwise free text.
bool next_line(std::istream& is, std::string& s) { return (bool)std::getline(is, s); }

double after_colon(const std::string& s) {
  size_t p = s.find(':');
  if (p == std::string::npos) return NAN;
  return std::strtod(s.c_str() + p + 1, nullptr);
}

bool read_inp(Case& c, const char* path) {
  std::ifstream f(path);
  if (!f) { c.err = std::string("cannot open ") + path; return false; }
  std::string s;
  next_line(f, s); next_line(f, s);                                     // BCH:544-545
  double h[18];
  for (int i = 0; i < 18; i++) {                                         // BCH:547-564
    if (!next_line(f, s)) { c.err = "truncated header"; return false; }
    h[i] = after_colon(s);
    if (std::isnan(h[i])) { c.err = "bad header line: " + s; return false; }
  }
  int eType = (int)h[0];
  c.NE = (int)h[1]; c.NCN = (int)h[2];
  if (eType != 1 || (int)h[3] != 27 || (int)h[4] != 8) { c.err = "only hex27/hex8 supported"; return false; }
  c.NGP = (int)h[5];
  c.dt = h[7]; c.t_ini = h[8]; c.t_final = h[9]; c.maxIter = (int)h[10];
  c.tolerance = h[11]; c.converge = h[12];
  c.density = h[14]; c.viscosity = h[15];
  next_line(f, s); next_line(f, s);                                     // BCH:579-580
  c.coord.assign((size_t)c.NCN * 3, 0.0);
  for (int i = 0; i < c.NCN; i++) {                                      // BCH:582-585
    if (!next_line(f, s)) { c.err = "truncated coordinates"; return false; }
    std::istringstream ls(s); int id;
    if (!(ls >> id >> c.coord[3*i] >> c.coord[3*i+1] >> c.coord[3*i+2])) { c.err = "bad coordinate line"; return false; }
  }
  next_line(f, s); next_line(f, s);                                     // BCH:614-615
  c.LtoGnode.assign((size_t)c.NE * 27, -1);
  for (int e = 0; e < c.NE; e++) {                                       // BCH:617-624
    if (!next_line(f, s)) { c.err = "truncated connectivity"; return false; }
    std::istringstream ls(s); int id; ls >> id;
    for (int i = 0; i < 8; i++) {
      int n; if (!(ls >> n)) { c.err = "bad connectivity line"; return false; }
      if (n < 1 || n > c.NCN) { c.err = "corner index out of range"; return false; }
      c.LtoGnode[(size_t)e*27 + i] = n - 1;
    }
  }
  next_line(f, s); next_line(f, s);                                     // BCH:628-629
  next_line(f, s); c.nBC = (int)after_colon(s);                         // BCH:631
  c.BCstr.assign((size_t)c.nBC * 3, 0.0);
  for (int i = 0; i < c.nBC; i++) {                                      // BCH:641-653
    next_line(f, s);
    size_t p = s.find(':');
    if (p == std::string::npos) { c.err = "bad BC line"; return false; }
    std::string rest = s.substr(p + 1);
    for (char& ch : rest) if (ch == ':') ch = ' ';
    std::istringstream ls(rest); double type;
    ls >> type >> c.BCstr[3*i] >> c.BCstr[3*i+1] >> c.BCstr[3*i+2];
  }
  next_line(f, s);
  next_line(f, s); int nVel = (int)after_colon(s);                      // BCH:655-656
  next_line(f, s); int nOut = (int)after_colon(s);                      // BCH:659-660
  next_line(f, s); next_line(f, s);                                     // BCH:666-667
  c.velFaces.assign((size_t)nVel * 3, 0);
  for (int i = 0; i < nVel; i++) {                                       // BCH:674-680
    if (!next_line(f, s)) { c.err = "truncated velocity BCs"; return false; }
    std::istringstream ls(s); int a, b, d;
    if (!(ls >> a >> b >> d)) { c.err = "bad velocity BC line"; return false; }
    c.velFaces[3*i] = a - 1; c.velFaces[3*i+1] = b - 1; c.velFaces[3*i+2] = d - 1;
  }
  next_line(f, s); next_line(f, s);                                     // BCH:684-685
  for (int i = 0; i < nOut; i++) next_line(f, s);                       // outflow faces are read but unused
  next_line(f, s); next_line(f, s);                                     // BCH:703-704
  next_line(f, s); c.zeroP = std::atoi(s.c_str()) - 1;                  // BCH:705-706
  next_line(f, s); next_line(f, s);                                     // BCH:712-713
  next_line(f, s);
  { std::istringstream ls(s); ls >> c.mon[0] >> c.mon[1] >> c.mon[2]; }
  c.NNp = c.NCN;                                                         // BCH:720-723
  return true;
}

// ---------------------------------------------------------------- topology
// BCH:732-783
bool findElemsOfPresNodes(Case& c) {
  const int LARGE = 10;
  c.elemsOfPres.assign((size_t)c.NNp * LARGE, -1);
  c.nElemOfPres.assign(c.NNp, 0);
  for (int e = 0; e < c.NE; e++)
    for (int i = 0; i < c.NENp; i++) {
      int node = c.LtoGnode[(size_t)e*27 + i];
      if (c.nElemOfPres[node] >= LARGE) { c.err = "more than 10 elements at a pressure node"; return false; }
      c.elemsOfPres[(size_t)node*LARGE + c.nElemOfPres[node]++] = e;
    }
  return true;
}

// BCH:790-846
void findElemNeighbors(Case& c) {
  const int LARGE = 26;
  c.neigh.assign((size_t)c.NE * LARGE, -1);
  c.nNeigh.assign(c.NE, 0);
  for (int e = 0; e < c.NE; e++) {
    for (int i = 0; i < c.NEC; i++) {
      int node = c.LtoGnode[(size_t)e*27 + i];
      for (int j = 0; j < c.nElemOfPres[node]; j++) {
        int el = c.elemsOfPres[(size_t)node*10 + j];
        if (el == e) continue;
        bool in = false;
        for (int k = 0; k < c.nNeigh[e]; k++) if (c.neigh[(size_t)e*LARGE + k] == el) { in = true; break; }
        if (!in) c.neigh[(size_t)e*LARGE + c.nNeigh[e]++] = el;
      }
    }
  }
}

// BCH:853-947, including the reference's `i = 0` reset inside the for loop
// (the loop increment then skips neighbour 0 after a colour bump).
bool setupMeshColoring(Case& c) {
  const int LARGE = 30;
  c.meshColors.assign(c.NE, -1);
  c.nMeshColors.assign(LARGE, 0);
  c.elementsOfColor.assign(c.NE, 0);
  for (int e = 0; e < c.NE; e++) {
    int cand = 0;
    for (int i = 0; i < c.nNeigh[e]; i++) {
      int el = c.neigh[(size_t)e*26 + i];
      if (cand == c.meshColors[el]) { cand++; i = 0; }
    }
    if (cand >= LARGE) { c.err = "more than 30 colours"; return false; }
    c.meshColors[e] = cand;
    c.nMeshColors[cand]++;
  }
  int count = 0;
  c.nActiveColors = 0;
  for (int i = 0; i < LARGE; i++) {
    if (c.nMeshColors[i] > 0) {
      c.nActiveColors++;
      for (int e = 0; e < c.NE; e++) if (c.meshColors[e] == i) c.elementsOfColor[count++] = e;
    }
  }
  return true;
}

// BCH:954-1320. Local node order: 0-7 corners, 8-19 edges, 20-25 faces, 26 centre.
const int kEdge[12][2] = {{0,1},{1,2},{2,3},{3,0},{0,4},{1,5},{2,6},{3,7},{4,5},{5,6},{6,7},{7,4}};
const int kFace[6][4] = {{0,1,2,3},{0,1,4,5},{1,2,5,6},{2,3,6,7},{0,3,4,7},{4,5,6,7}};
// Non-corner nodes of each face, BCH:1480-1527
const int kFaceMid[6][5] = {{8,9,10,11,20},{8,12,13,16,21},{9,13,14,17,22},{10,14,15,18,23},{11,12,15,19,24},{16,17,18,19,25}};

void setupNonCornerNodes(Case& c) {
  const double SMALL = 1e-10;
  const int NE = c.NE;
  std::vector<double> xyz((size_t)NE * 27 * 3 + (size_t)c.NCN * 3);  // generous, like NE*NENv in BCH:570
  std::copy(c.coord.begin(), c.coord.end(), xyz.begin());
  int nodeCount = c.NCN;
  auto L = [&](int e, int i) -> int& { return c.LtoGnode[(size_t)e*27 + i]; };
  std::vector<int> nb;
  nb.reserve(26 * 12);
  for (int e = 0; e < NE; e++) {                                        // BCH:977-1134
    nb.clear();
    for (int e2 = 0; e2 < c.nNeigh[e]; e2++) {
      int ne = c.neigh[(size_t)e*26 + e2];
      for (int i = c.NEC; i < c.NEC + c.NEE; i++) if (L(ne, i) != -1) nb.push_back(L(ne, i));
    }
    for (int ed = 0; ed < 12; ed++) {
      int n1 = L(e, kEdge[ed][0]), n2 = L(e, kEdge[ed][1]);
      double m[3];
      for (int d = 0; d < 3; d++) m[d] = 0.5 * (xyz[3*(size_t)n1+d] + xyz[3*(size_t)n2+d]);
      bool found = false;
      for (int node : nb) {
        if (std::fabs(m[0] - xyz[3*(size_t)node]) < SMALL && std::fabs(m[1] - xyz[3*(size_t)node+1]) < SMALL &&
            std::fabs(m[2] - xyz[3*(size_t)node+2]) < SMALL) { L(e, ed + c.NEC) = node; found = true; break; }
      }
      if (!found) {
        L(e, ed + c.NEC) = nodeCount;
        for (int d = 0; d < 3; d++) xyz[3*(size_t)nodeCount+d] = m[d];
        nodeCount++;
      }
    }
  }
  for (int e = 0; e < NE; e++) {                                        // BCH:1140-1228
    nb.clear();
    for (int e2 = 0; e2 < c.nNeigh[e]; e2++) {
      int ne = c.neigh[(size_t)e*26 + e2];
      for (int i = c.NEC + c.NEE; i < c.NENv; i++) if (L(ne, i) != -1) nb.push_back(L(ne, i));
    }
    for (int f = 0; f < 6; f++) {
      int n[4];
      for (int k = 0; k < 4; k++) n[k] = L(e, kFace[f][k]);
      double m[3];
      for (int d = 0; d < 3; d++)
        m[d] = 0.25 * (xyz[3*(size_t)n[0]+d] + xyz[3*(size_t)n[1]+d] + xyz[3*(size_t)n[2]+d] + xyz[3*(size_t)n[3]+d]);
      bool found = false;
      for (int node : nb) {
        if (std::fabs(m[0] - xyz[3*(size_t)node]) < SMALL && std::fabs(m[1] - xyz[3*(size_t)node+1]) < SMALL &&
            std::fabs(m[2] - xyz[3*(size_t)node+2]) < SMALL) { L(e, f + c.NEC + c.NEE) = node; found = true; break; }
      }
      if (!found) {
        L(e, f + c.NEC + c.NEE) = nodeCount;
        for (int d = 0; d < 3; d++) xyz[3*(size_t)nodeCount+d] = m[d];
        nodeCount++;
      }
    }
  }
  for (int e = 0; e < NE; e++) {                                        // BCH:1235-1259
    double m[3];
    for (int d = 0; d < 3; d++) {
      double s = 0.0;
      for (int k = 0; k < 8; k++) s += xyz[3*(size_t)L(e, k)+d];
      m[d] = 0.125 * s;
    }
    L(e, 26) = nodeCount;
    for (int d = 0; d < 3; d++) xyz[3*(size_t)nodeCount+d] = m[d];
    nodeCount++;
  }
  c.NN = nodeCount;
  xyz.resize((size_t)c.NN * 3);
  c.coord.swap(xyz);
}

// BCH:1410-1580: last writer wins; nodes listed in increasing node order.
void determineVelBCnodes(Case& c) {
  std::vector<int> info(c.NN, -1);
  int nVel = (int)c.velFaces.size() / 3;
  for (int i = 0; i < nVel; i++) {
    int e = c.velFaces[3*i], f = c.velFaces[3*i+1], w = c.velFaces[3*i+2];
    for (int k = 0; k < 4; k++) info[c.LtoGnode[(size_t)e*27 + kFace[f][k]]] = w;
    for (int k = 0; k < 5; k++) info[c.LtoGnode[(size_t)e*27 + kFaceMid[f][k]]] = w;
  }
  c.BCnodes.clear(); c.BCwhich.clear();
  for (int i = 0; i < c.NN; i++) if (info[i] != -1) { c.BCnodes.push_back(i); c.BCwhich.push_back(info[i]); }
}

// BCH:1587-1637
bool findElemsOfVelNodes(Case& c) {
  const int LARGE = 10;
  c.elemsOfVel.assign((size_t)c.NN * LARGE, -1);
  c.nElemOfVel.assign(c.NN, 0);
  for (int e = 0; e < c.NE; e++)
    for (int i = 0; i < 27; i++) {
      int node = c.LtoGnode[(size_t)e*27 + i];
      if (c.nElemOfVel[node] >= LARGE) { c.err = "more than 10 elements at a velocity node"; return false; }
      c.elemsOfVel[(size_t)node*LARGE + c.nElemOfVel[node]++] = e;
    }
  return true;
}

// BCH:1644-1668 (corner nodes only, first strict minimum wins)
void findMonitorPoint(Case& c) {
  double dist = 1e6;
  for (int i = 0; i < c.NCN; i++) {
    double dx = c.coord[3*(size_t)i] - c.mon[0], dy = c.coord[3*(size_t)i+1] - c.mon[1], dz = c.coord[3*(size_t)i+2] - c.mon[2];
    double r = std::sqrt(dx*dx + dy*dy + dz*dz);
    if (r < dist) { dist = r; c.monPoint = i; }
  }
}

// BCH:1675-1921 (M/K/A pattern) and BCH:1928-2159 (G pattern). Columns of a
// row are the union of the nodes of the elements at the row's node, sorted.
void buildPattern(const Case& c, int ncolsPerElem, bool pres, std::vector<int>& rowptr, std::vector<int>& col) {
  const int NN = c.NN;
  std::vector<std::vector<int>> rows(NN);
  #pragma omp parallel for schedule(dynamic, 4096)
  for (int r = 0; r < NN; r++) {
    std::vector<int>& v = rows[r];
    for (int i = 0; i < c.nElemOfVel[r]; i++) {
      int e = c.elemsOfVel[(size_t)r*10 + i];
      for (int j = 0; j < ncolsPerElem; j++) v.push_back(c.LtoGnode[(size_t)e*27 + j]);
    }
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
  }
  (void)pres;
  rowptr.assign(NN + 1, 0);
  for (int r = 0; r < NN; r++) rowptr[r+1] = rowptr[r] + (int)rows[r].size();
  col.resize(rowptr[NN]);
  #pragma omp parallel for schedule(dynamic, 4096)
  for (int r = 0; r < NN; r++) std::copy(rows[r].begin(), rows[r].end(), col.begin() + rowptr[r]);
}

int findInRow(const std::vector<int>& rowptr, const std::vector<int>& col, int r, int c) {
  auto b = col.begin() + rowptr[r], e = col.begin() + rowptr[r+1];
  auto it = std::lower_bound(b, e, c);
  return (int)(it - col.begin());   // the element is always present
}

void setupSparse(Case& c) {
  buildPattern(c, 27, false, c.Mrow, c.Mcol);
  buildPattern(c, 8, true, c.Grow, c.Gcol);
  c.mapM.resize((size_t)c.NE * 729);
  c.mapG.resize((size_t)c.NE * 216);
  #pragma omp parallel for schedule(static)
  for (int e = 0; e < c.NE; e++) {                                      // BCH:1861-1898, 2103-2125
    const int* L = &c.LtoGnode[(size_t)e*27];
    for (int i = 0; i < 27; i++) {
      for (int j = 0; j < 27; j++) c.mapM[(size_t)e*729 + i*27 + j] = findInRow(c.Mrow, c.Mcol, L[i], L[j]);
      for (int j = 0; j < 8; j++) c.mapG[(size_t)e*216 + i*8 + j] = findInRow(c.Grow, c.Gcol, L[i], L[j]);
    }
  }
}

// ---------------------------------------------------------------- reference element
// BCH:2166-2208; NGP=27 is the tensor 3-point Gauss-Legendre rule (points
// 0, +-sqrt(3/5), weights 8/9, 5/9) with xi running fastest, like NGP=8.
bool setupGQ(Case& c) {
  c.gqp.assign((size_t)c.NGP * 3, 0.0);
  c.gqw.assign(c.NGP, 0.0);
  if (c.NGP == 1) { c.gqw[0] = 4.0; return true; }                     // BCH:2180-2182 (sic)
  if (c.NGP == 8) {
    const double a = std::sqrt(1. / 3);
    for (int k = 0; k < 8; k++) {
      c.gqp[3*k] = (k & 1) ? a : -a; c.gqp[3*k+1] = (k & 2) ? a : -a; c.gqp[3*k+2] = (k & 4) ? a : -a;
      c.gqw[k] = 1.0;
    }
    return true;
  }
  if (c.NGP == 27) {
    const double p[3] = {-std::sqrt(0.6), 0.0, std::sqrt(0.6)};
    const double w[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
    for (int kz = 0; kz < 3; kz++) for (int ky = 0; ky < 3; ky++) for (int kx = 0; kx < 3; kx++) {
      int k = kx + 3*ky + 9*kz;
      c.gqp[3*k] = p[kx]; c.gqp[3*k+1] = p[ky]; c.gqp[3*k+2] = p[kz];
      c.gqw[k] = w[kx] * w[ky] * w[kz];
    }
    return true;
  }
  c.err = "NGP must be 1, 8 or 27";
  return false;
}

// Parametric position (-1, 0, +1 per axis) of each hex27 local node, BCH:2312-2343.
const int kNodePos[27][3] = {
  {-1,-1,-1},{1,-1,-1},{1,1,-1},{-1,1,-1},{-1,-1,1},{1,-1,1},{1,1,1},{-1,1,1},
  {0,-1,-1},{1,0,-1},{0,1,-1},{-1,0,-1},
  {-1,-1,0},{1,-1,0},{1,1,0},{-1,1,0},
  {0,-1,1},{1,0,1},{0,1,1},{-1,0,1},
  {0,0,-1},{0,-1,0},{1,0,0},{0,1,0},{-1,0,0},{0,0,1},{0,0,0}};

// 1-D quadratic Lagrange factor used by BCH:2312-2450: (x^2-x)/2, 1-x^2, (x^2+x)/2
inline double q1(int a, double x) { return a < 0 ? 0.5 * (x*x - x) : (a > 0 ? 0.5 * (x*x + x) : (1 - x*x)); }
inline double dq1(int a, double x) { return a < 0 ? 0.5 * (2*x - 1) : (a > 0 ? 0.5 * (2*x + 1) : (-2*x)); }

void calcShape(Case& c) {                                                // BCH:2215-2488
  const int G = c.NGP;
  c.Sv.assign((size_t)G * 27, 0.0); c.Sp.assign((size_t)G * 8, 0.0);
  c.dSv.assign((size_t)3 * 27 * G, 0.0); c.dSp.assign((size_t)3 * 8 * G, 0.0);
  for (int k = 0; k < G; k++) {
    double x = c.gqp[3*k], y = c.gqp[3*k+1], z = c.gqp[3*k+2];
    for (int i = 0; i < 8; i++) {
      double sx = kNodePos[i][0], sy = kNodePos[i][1], sz = kNodePos[i][2];
      c.Sp[k*8 + i] = 0.125 * (1 + sx*x) * (1 + sy*y) * (1 + sz*z);
      c.dSp[(0*8 + i)*G + k] = 0.125 * sx * (1 + sy*y) * (1 + sz*z);
      c.dSp[(1*8 + i)*G + k] = 0.125 * (1 + sx*x) * sy * (1 + sz*z);
      c.dSp[(2*8 + i)*G + k] = 0.125 * (1 + sx*x) * (1 + sy*y) * sz;
    }
    for (int i = 0; i < 27; i++) {
      int a = kNodePos[i][0], b = kNodePos[i][1], d = kNodePos[i][2];
      c.Sv[k*27 + i] = q1(a, x) * q1(b, y) * q1(d, z);
      c.dSv[(0*27 + i)*G + k] = dq1(a, x) * q1(b, y) * q1(d, z);
      c.dSv[(1*27 + i)*G + k] = q1(a, x) * dq1(b, y) * q1(d, z);
      c.dSv[(2*27 + i)*G + k] = q1(a, x) * q1(b, y) * dq1(d, z);
    }
  }
}

// BCH:2538-2635: J = dSp * x_corners, cofactor inverse, gDSv = J^-1 dSv.
// Fills gdsv[(k*27 + j)*3 + d] and fac[k] = detJ * w_k for one element.
void elementGeometry(const Case& c, int e, double* gdsv, double* fac) {
  const int G = c.NGP;
  double ec[8][3];
  for (int i = 0; i < 8; i++) {
    int n = c.LtoGnode[(size_t)e*27 + i];
    for (int d = 0; d < 3; d++) ec[i][d] = c.coord[3*(size_t)n + d];
  }
  for (int k = 0; k < G; k++) {
    double J[3][3], iJ[3][3];
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) {
      double s = 0;
      for (int m = 0; m < 8; m++) s = s + c.dSp[(i*8 + m)*G + k] * ec[m][j];
      J[i][j] = s;
    }
    iJ[0][0] =   J[1][1]*J[2][2] - J[2][1]*J[1][2];
    iJ[0][1] = -(J[0][1]*J[2][2] - J[0][2]*J[2][1]);
    iJ[0][2] =   J[0][1]*J[1][2] - J[1][1]*J[0][2];
    iJ[1][0] = -(J[1][0]*J[2][2] - J[1][2]*J[2][0]);
    iJ[1][1] =   J[2][2]*J[0][0] - J[2][0]*J[0][2];
    iJ[1][2] = -(J[1][2]*J[0][0] - J[1][0]*J[0][2]);
    iJ[2][0] =   J[1][0]*J[2][1] - J[2][0]*J[1][1];
    iJ[2][1] = -(J[2][1]*J[0][0] - J[2][0]*J[0][1]);
    iJ[2][2] =   J[1][1]*J[0][0] - J[1][0]*J[0][1];
    double det = J[0][0]*(J[1][1]*J[2][2] - J[2][1]*J[1][2]) +
                 J[0][1]*(J[1][2]*J[2][0] - J[1][0]*J[2][2]) +
                 J[0][2]*(J[1][0]*J[2][1] - J[1][1]*J[2][0]);
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) iJ[i][j] = iJ[i][j] / det;
    for (int i = 0; i < 3; i++) for (int j = 0; j < 27; j++) {
      double s = 0;
      for (int m = 0; m < 3; m++) s = s + iJ[i][m] * c.dSv[(m*27 + j)*G + k];
      gdsv[(k*27 + j)*3 + i] = s;
    }
    fac[k] = det * c.gqw[k];                                             // BCH:2644-2651
  }
}

// ---------------------------------------------------------------- step0
// BCH:3127-3315
void step0(Case& c) {
  const int G = c.NGP;
  size_t nnzM = c.Mcol.size(), nnzG = c.Gcol.size();
  c.Mval.assign(nnzM, 0.0); c.Kval.assign(nnzM, 0.0);
  c.G1.assign(nnzG, 0.0); c.G2.assign(nnzG, 0.0); c.G3.assign(nnzG, 0.0);
  const double invRho = 1.0 / c.density;
  const int NE = c.NE;
  // element matrices in parallel, scatter sequentially in element order
  const int CH = 2048;
  std::vector<double> buf((size_t)CH * (729*2 + 216*3));
  for (int e0 = 0; e0 < NE; e0 += CH) {
    int e1 = std::min(NE, e0 + CH);
    #pragma omp parallel for schedule(static)
    for (int e = e0; e < e1; e++) {
      double* Me = &buf[(size_t)(e - e0) * (729*2 + 648)];
      double* Ke = Me + 729;
      double* Ge1 = Ke + 729; double* Ge2 = Ge1 + 216; double* Ge3 = Ge2 + 216;
      std::vector<double> gd((size_t)G * 81), fac(G);
      elementGeometry(c, e, gd.data(), fac.data());
      for (int i = 0; i < 729; i++) { Me[i] = 0.0; Ke[i] = 0.0; }
      for (int i = 0; i < 216; i++) { Ge1[i] = 0.0; Ge2[i] = 0.0; Ge3[i] = 0.0; }
      for (int k = 0; k < G; k++) {                                      // BCH:3176-3231
        double f = fac[k];
        const double* g = &gd[(size_t)k * 81];
        for (int i = 0; i < 27; i++) for (int j = 0; j < 27; j++) {
          Me[i*27 + j] = Me[i*27 + j] + c.Sv[k*27 + i] * c.Sv[k*27 + j] * f;
          Ke[i*27 + j] = Ke[i*27 + j] + c.viscosity * (g[i*3]*g[j*3] + g[i*3+1]*g[j*3+1] + g[i*3+2]*g[j*3+2]) * f;
        }
        for (int i = 0; i < 27; i++) for (int j = 0; j < 8; j++) {
          Ge1[i*8 + j] = Ge1[i*8 + j] - invRho * c.Sp[k*8 + j] * g[i*3] * f;
          Ge2[i*8 + j] = Ge2[i*8 + j] - invRho * c.Sp[k*8 + j] * g[i*3+1] * f;
          Ge3[i*8 + j] = Ge3[i*8 + j] - invRho * c.Sp[k*8 + j] * g[i*3+2] * f;
        }
      }
    }
    for (int e = e0; e < e1; e++) {                                      // BCH:3235-3247
      const double* Me = &buf[(size_t)(e - e0) * (729*2 + 648)];
      const double* Ke = Me + 729;
      const double* Ge1 = Ke + 729; const double* Ge2 = Ge1 + 216; const double* Ge3 = Ge2 + 216;
      const int* mM = &c.mapM[(size_t)e*729];
      const int* mG = &c.mapG[(size_t)e*216];
      for (int i = 0; i < 729; i++) { c.Mval[mM[i]] += Me[i]; c.Kval[mM[i]] += Ke[i]; }
      for (int i = 0; i < 216; i++) { c.G1[mG[i]] += Ge1[i]; c.G2[mG[i]] += Ge2[i]; c.G3[mG[i]] += Ge3[i]; }
    }
  }
  const int NN = c.NN;
  c.Md.assign((size_t)3*NN, 0.0);
  for (int r = 0; r < NN; r++)                                           // BCH:3262-3266
    for (int p = c.Mrow[r]; p < c.Mrow[r+1]; p++) c.Md[r] += c.Mval[p];
  for (int i = 0; i < NN; i++) { c.Md[i + NN] = c.Md[i]; c.Md[i + 2*NN] = c.Md[i]; }
  c.MdOrig = c.Md;
  c.MdOrigInv.resize((size_t)3*NN);
  for (int i = 0; i < 3*NN; i++) c.MdOrigInv[i] = 1.0 / c.MdOrig[i];   // BCH:3287-3289
  for (int n : c.BCnodes) { c.Md[n] = 1.0; c.Md[n + NN] = 1.0; c.Md[n + 2*NN] = 1.0; }  // BCH:4121-4135
  c.MdInv.resize((size_t)3*NN);
  for (int i = 0; i < 3*NN; i++) c.MdInv[i] = 1.0 / c.Md[i];           // BCH:3292-3295
}

// BCH:3322-3479 + BCH:4154-4186. Z(p,q) = sum_c sum_{i ascending} G_c(i,p) *
// (G_c(i,q) * MdOrigInv[i]), the three products summed as (Z1 + Z2) + Z3,
// which is the value order of three cs_multiply calls followed by two
// value-wise adds. The pattern is structural (numerical zeros kept).
void buildGT(Case& c) {
  const int NN = c.NN, NNp = c.NNp;
  c.Trow.assign(NNp + 1, 0); c.Tcol.resize(c.Gcol.size()); c.Tpos.resize(c.Gcol.size());
  for (int q : c.Gcol) c.Trow[q + 1]++;
  for (int p = 0; p < NNp; p++) c.Trow[p + 1] += c.Trow[p];
  std::vector<int> next(c.Trow.begin(), c.Trow.end() - 1);
  for (int i = 0; i < NN; i++)
    for (int k = c.Grow[i]; k < c.Grow[i+1]; k++) { int p = c.Gcol[k]; c.Tcol[next[p]] = i; c.Tpos[next[p]] = k; next[p]++; }
}

void calculateZ(Case& c) {
  const int NNp = c.NNp;
  buildGT(c);
  const std::vector<int>& Trow = c.Trow; const std::vector<int>& Tcol = c.Tcol; const std::vector<int>& Tpos = c.Tpos;
  std::vector<std::vector<int>> zc(NNp);
  std::vector<std::vector<double>> zv(NNp);
  #pragma omp parallel
  {
    std::vector<int> mark(NNp, -1), slot(NNp, 0);
    #pragma omp for schedule(dynamic, 256)
    for (int p = 0; p < NNp; p++) {
      std::vector<int>& cols = zc[p];
      for (int t = Trow[p]; t < Trow[p+1]; t++) {
        int i = Tcol[t];
        for (int k = c.Grow[i]; k < c.Grow[i+1]; k++) {
          int q = c.Gcol[k];
          if (mark[q] != p) { mark[q] = p; cols.push_back(q); }
        }
      }
      std::sort(cols.begin(), cols.end());
      for (size_t s = 0; s < cols.size(); s++) slot[cols[s]] = (int)s;
      std::vector<double> z1(cols.size(), 0.0), z2(cols.size(), 0.0), z3(cols.size(), 0.0);
      for (int t = Trow[p]; t < Trow[p+1]; t++) {
        int i = Tcol[t], kp = Tpos[t];
        double m = c.MdOrigInv[i];
        double a1 = c.G1[kp], a2 = c.G2[kp], a3 = c.G3[kp];
        for (int k = c.Grow[i]; k < c.Grow[i+1]; k++) {
          int s = slot[c.Gcol[k]];
          z1[s] += a1 * (c.G1[k] * m);
          z2[s] += a2 * (c.G2[k] * m);
          z3[s] += a3 * (c.G3[k] * m);
        }
      }
      std::vector<double>& v = zv[p];
      v.resize(cols.size());
      for (size_t s = 0; s < cols.size(); s++) v[s] = (z1[s] + z2[s]) + z3[s];
    }
  }
  c.Zrow.assign(NNp + 1, 0);
  for (int p = 0; p < NNp; p++) c.Zrow[p+1] = c.Zrow[p] + (int)zc[p].size();
  c.Zcol.resize(c.Zrow[NNp]); c.Zval.resize(c.Zrow[NNp]);
  for (int p = 0; p < NNp; p++) {
    std::copy(zc[p].begin(), zc[p].end(), c.Zcol.begin() + c.Zrow[p]);
    std::copy(zv[p].begin(), zv[p].end(), c.Zval.begin() + c.Zrow[p]);
  }
  const double LARGE = 1000;                                             // BCH:4167-4177 (Q5: node > 0)
  int node = c.zeroP;
  if (node > 0)
    for (int k = c.Zrow[node]; k < c.Zrow[node+1]; k++) if (c.Zcol[k] == node) { c.Zval[k] = c.Zval[k] * LARGE; break; }
}

// ---------------------------------------------------------------- sparse kernels
void csrmv(const std::vector<int>& rp, const std::vector<int>& ci, const double* v, const double* x, double* y,
           int nrows, double alpha, double beta) {
  #pragma omp parallel for schedule(static)
  for (int r = 0; r < nrows; r++) {
    double s = 0.0;
    for (int k = rp[r]; k < rp[r+1]; k++) s += v[k] * x[ci[k]];
    y[r] = (beta == 0.0 ? 0.0 : beta * y[r]) + alpha * s;
  }
}

// y = alpha * G_d^T x + beta*y, gathered over the transposed pattern so each
// output sums its terms in ascending velocity row, the order of a serial
// row-by-row scatter (the MKL 't' csrmv of BCH:3864-3868).
void gtmv(const Case& c, const double* v, const double* x, double* y, double alpha, double beta) {
  #pragma omp parallel for schedule(static)
  for (int p = 0; p < c.NNp; p++) {
    double s = (beta == 0.0) ? 0.0 : beta * y[p];
    for (int t = c.Trow[p]; t < c.Trow[p+1]; t++) s += v[c.Tpos[t]] * (alpha * x[c.Tcol[t]]);
    y[p] = s;
  }
}

// Textbook Jacobi-PCG (BCH:3981-4055 restated): x0 = 0, stop when
// ||r||_2 <= rtol * ||b||_2 or after maxit iterations. With warm, x0 is the
// x passed in (the reference's GPU path keeps the previous Pdot, CU:425-442)
// and r0 = b - A x0; the stop test stays relative to ||b||.
int pcg(const std::vector<int>& rp, const std::vector<int>& ci, const std::vector<double>& v, int n,
        const double* b, double* x, double rtol, int maxit, double* relres, bool warm = false) {
  std::vector<double> dinv(n), r(n), z(n), p(n), q(n);
  for (int i = 0; i < n; i++) {
    double d = 0.0;
    for (int k = rp[i]; k < rp[i+1]; k++) if (ci[k] == i) d = v[k];
    dinv[i] = 1.0 / d;
  }
  double bb = 0.0;
  if (warm) {
    csrmv(rp, ci, v.data(), x, r.data(), n, 1.0, 0.0);
    for (int i = 0; i < n; i++) { r[i] = b[i] - r[i]; bb += b[i] * b[i]; }
  } else {
    for (int i = 0; i < n; i++) { x[i] = 0.0; r[i] = b[i]; bb += b[i] * b[i]; }
  }
  double bnorm = std::sqrt(bb);
  if (bnorm == 0.0) { *relres = 0.0; return 0; }
  double rr0 = 0.0;
  for (int i = 0; i < n; i++) rr0 += r[i] * r[i];
  if (std::sqrt(rr0) <= rtol * bnorm) { *relres = std::sqrt(rr0) / bnorm; return 0; }
  double rz = 0.0;
  for (int i = 0; i < n; i++) { z[i] = dinv[i] * r[i]; p[i] = z[i]; rz += r[i] * z[i]; }
  int it = 0;
  double rnorm = std::sqrt(rr0);
  while (it < maxit) {
    csrmv(rp, ci, v.data(), p.data(), q.data(), n, 1.0, 0.0);
    double pq = 0.0;
    #pragma omp parallel for reduction(+:pq)
    for (int i = 0; i < n; i++) pq += p[i] * q[i];
    double alpha = rz / pq;
    double rr = 0.0;
    #pragma omp parallel for reduction(+:rr)
    for (int i = 0; i < n; i++) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; rr += r[i] * r[i]; }
    it++;
    rnorm = std::sqrt(rr);
    if (rnorm <= rtol * bnorm) break;
    double rzn = 0.0;
    #pragma omp parallel for reduction(+:rzn)
    for (int i = 0; i < n; i++) { z[i] = dinv[i] * r[i]; rzn += r[i] * z[i]; }
    double beta = rzn / rz;
    rz = rzn;
    #pragma omp parallel for
    for (int i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
  }
  *relres = rnorm / bnorm;
  return it;
}

// ---------------------------------------------------------------- time stepping
// BCH:2718-2808 (allocation, IC), BCH:4062-4107 (applyBC_initial)
void initialize(Case& c) {
  const size_t n3 = (size_t)3 * c.NN, np = c.NNp;
  for (auto* v : {&c.Un, &c.Unp1, &c.Unp1_prev, &c.UnpHalf, &c.UnpHalf_prev, &c.Acc, &c.Acc_prev, &c.KAcc, &c.R1, &c.R3})
    v->assign(n3, 0.0);
  for (auto* v : {&c.Pn, &c.Pnp1, &c.Pnp1_prev, &c.Pdot, &c.R2}) v->assign(np, 0.0);
  for (size_t i = 0; i < c.BCnodes.size(); i++) {
    int node = c.BCnodes[i], w = c.BCwhich[i];
    c.Un[node] = c.BCstr[3*w]; c.Un[node + c.NN] = c.BCstr[3*w+1]; c.Un[node + 2*(size_t)c.NN] = c.BCstr[3*w+2];
  }
  c.timeN = 0;
  c.timeT = c.t_ini;
}

// BCH:3520-3685. R1 <- -A(Un) * UnpHalf_prev, per colour; the element
// contributions are computed in parallel and scattered in colour order.
void convection(const Case& c, const double* Un, const double* Uprev, double* R1) {
  const int NN = c.NN, G = c.NGP, NE = c.NE;
  for (int i = 0; i < 3*NN; i++) R1[i] = 0.0;                          // BCH:3541-3545
  std::vector<double> r1e((size_t)NE * 81);
  #pragma omp parallel
  {
    std::vector<double> gd((size_t)G * 81), fac(G), Ae(729);
    #pragma omp for schedule(static)
    for (int e = 0; e < NE; e++) {
      const int* L = &c.LtoGnode[(size_t)e*27];
      double u0n[27], v0n[27], w0n[27], up[27], vp[27], wp[27];
      for (int i = 0; i < 27; i++) {
        int g = L[i];
        u0n[i] = Un[g]; up[i] = Uprev[g];
        v0n[i] = Un[g + NN]; vp[i] = Uprev[g + NN];
        w0n[i] = Un[g + 2*NN]; wp[i] = Uprev[g + 2*NN];
      }
      elementGeometry(c, e, gd.data(), fac.data());
      std::fill(Ae.begin(), Ae.end(), 0.0);
      for (int k = 0; k < G; k++) {
        double f = fac[k];
        double u0 = 0, v0 = 0, w0 = 0;
        for (int i = 0; i < 27; i++) {
          u0 = u0 + c.Sv[k*27 + i] * u0n[i]; v0 = v0 + c.Sv[k*27 + i] * v0n[i]; w0 = w0 + c.Sv[k*27 + i] * w0n[i];
        }
        const double* g = &gd[(size_t)k * 81];
        for (int i = 0; i < 27; i++) for (int j = 0; j < 27; j++)
          Ae[i*27 + j] = Ae[i*27 + j] + (u0 * g[j*3] + v0 * g[j*3+1] + w0 * g[j*3+2]) * c.Sv[k*27 + i] * f;
      }
      double* o = &r1e[(size_t)e * 81];
      for (int i = 0; i < 27; i++) {
        double a = 0, b = 0, d = 0;
        for (int j = 0; j < 27; j++) { a += Ae[i*27 + j] * up[j]; b += Ae[i*27 + j] * vp[j]; d += Ae[i*27 + j] * wp[j]; }
        o[i] = a; o[27 + i] = b; o[54 + i] = d;
      }
    }
  }
  for (int t = 0; t < NE; t++) {
    int e = c.elementsOfColor[t];
    const int* L = &c.LtoGnode[(size_t)e*27];
    const double* o = &r1e[(size_t)e * 81];
    for (int i = 0; i < 27; i++) {
      R1[L[i]] -= o[i]; R1[L[i] + NN] -= o[27 + i]; R1[L[i] + 2*NN] -= o[54 + i];
    }
  }
}

// Consistent-mass velocity update (no counterpart in the reference, which
// lumps M, BCH:3262-3295): a_d = M_bc^-1 R_d for d = x, y, z with M_bc = M
// whose BC rows and columns are replaced by identity rows and columns. R is
// zero at BC nodes, so a is too, as with the lumped update (MdInv = 1 there).
// Each component is an independent textbook Jacobi-PCG to mass_rtol.
void mass_solve(Case& c, const double* R, double* a) {
  const int NN = c.NN;
  if (c.Mbc.empty()) {
    std::vector<char> bc(NN, 0);
    for (int n : c.BCnodes) bc[n] = 1;
    c.Mbc = c.Mval;
    for (int r = 0; r < NN; r++)
      for (int k = c.Mrow[r]; k < c.Mrow[r+1]; k++)
        if (bc[r] || bc[c.Mcol[k]]) c.Mbc[k] = (c.Mcol[k] == r) ? 1.0 : 0.0;
  }
  c.last_mass_iters = 0;
  c.last_mass_relres = 0.0;
  for (int d = 0; d < 3; d++) {
    double rel = 0.0;
    const int it = pcg(c.Mrow, c.Mcol, c.Mbc, NN, R + (size_t)d*NN, a + (size_t)d*NN, c.mass_rtol, c.mass_maxit, &rel);
    c.last_mass_iters = std::max(c.last_mass_iters, it);
    c.last_mass_relres = std::max(c.last_mass_relres, rel);
    c.total_mass_iters += it;
  }
}

// the operand of a K product: v itself, or v rounded to FP32 (f32_operands)
std::vector<double> k_operand(const Case& c, const std::vector<double>& v) {
  if (!c.f32_operands || c.mass_mode != 0) return v;
  std::vector<double> r(v.size());
  for (size_t i = 0; i < v.size(); i++) r[i] = (double)(float)v[i];
  return r;
}

// BCH:3692-3801
void step1(Case& c, int iter) {
  (void)iter;
  const int NN = c.NN;
  convection(c, c.Un.data(), c.UnpHalf_prev.data(), c.R1.data());
  const std::vector<double> uk = k_operand(c, c.UnpHalf_prev);
  for (int d = 0; d < 3; d++) {
    csrmv(c.Mrow, c.Mcol, c.Kval.data(), uk.data() + (size_t)d*NN, c.R1.data() + (size_t)d*NN, NN, -1.0, 1.0);
  }
  const std::vector<double>* Gs[3] = {&c.G1, &c.G2, &c.G3};
  for (int d = 0; d < 3; d++)
    csrmv(c.Grow, c.Gcol, Gs[d]->data(), c.Pn.data(), c.R1.data() + (size_t)d*NN, NN, -1.0, 1.0);
  for (int n : c.BCnodes) { c.R1[n] = 0.0; c.R1[n + NN] = 0.0; c.R1[n + 2*(size_t)NN] = 0.0; }   // BCH:4137-4146
  if (c.mass_mode == 1) {
    std::vector<double> a((size_t)3*NN);
    mass_solve(c, c.R1.data(), a.data());
    for (size_t i = 0; i < (size_t)3*NN; i++) c.UnpHalf[i] = c.Un[i] + c.dt * a[i];
    return;
  }
  for (size_t i = 0; i < (size_t)3*NN; i++) c.UnpHalf[i] = c.Un[i] + c.dt * c.R1[i] * c.MdInv[i];
}

// Cholesky of Z in its own (natural) order, Z = L L^T with L stored as a
// lower band of half-bandwidth w = max |i - j| over Z's pattern (a band
// factor keeps no fill outside the band). The prototype factors Z with an
// AMD ordering (ML:1876-1885); the order changes the fill, not the solution.
// Returns false when a pivot is not positive.
bool chol_factor(Case& c) {
  const int n = c.NNp;
  int w = 0;
  for (int i = 0; i < n; i++)
    for (int k = c.Zrow[i]; k < c.Zrow[i+1]; k++) w = std::max(w, std::abs(i - c.Zcol[k]));
  const size_t W = (size_t)w + 1;
  std::vector<double> L((size_t)n * W, 0.0);
  for (int i = 0; i < n; i++)
    for (int k = c.Zrow[i]; k < c.Zrow[i+1]; k++)
      if (c.Zcol[k] <= i) L[i * W + (c.Zcol[k] - i + w)] = c.Zval[k];
  for (int j = 0; j < n; j++) {                          // left-looking, column j
    double* Lj = L.data() + j * W;
    double s = Lj[w];
    for (int k = std::max(0, j - w); k < j; k++) s -= Lj[k - j + w] * Lj[k - j + w];
    if (!(s > 0.0)) { c.chol_w = -1; return false; }
    const double d = std::sqrt(s);
    Lj[w] = d;
    const int iend = std::min(n - 1, j + w);
#pragma omp parallel for schedule(static)
    for (int i = j + 1; i <= iend; i++) {
      double* Li = L.data() + i * W;
      double t = Li[j - i + w];
      for (int k = std::max(0, i - w); k < j; k++) t -= Li[k - i + w] * Lj[k - j + w];
      Li[j - i + w] = t / d;
    }
  }
  c.chol_L.swap(L);
  c.chol_w = w;
  return true;
}

// x = Z^-1 b with the kept factor: L y = b, then L^T x = y (ML:2021-2030)
void chol_solve(const Case& c, const double* b, double* x) {
  const int n = c.NNp, w = c.chol_w;
  const size_t W = (size_t)w + 1;
  const double* L = c.chol_L.data();
  std::vector<double> y(n);
  for (int i = 0; i < n; i++) {
    double t = b[i];
    for (int k = std::max(0, i - w); k < i; k++) t -= L[i * W + (k - i + w)] * y[k];
    y[i] = t / L[i * W + w];
  }
  for (int i = n - 1; i >= 0; i--) {
    double t = y[i];
    for (int k = i + 1; k <= std::min(n - 1, i + w); k++) t -= L[k * W + (i - k + w)] * x[k];
    x[i] = t / L[i * W + w];
  }
}

// the direct pressure solve: factor on first use, solve, |b - Z x| / |b|
int direct_solve(Case& c, const double* b, double* x, double* relres) {
  if (c.chol_w < 0 && !chol_factor(c)) { c.err = "Cholesky of Z met a non-positive pivot"; return 1; }
  chol_solve(c, b, x);
  const int n = c.NNp;
  std::vector<double> r(n);
  csrmv(c.Zrow, c.Zcol, c.Zval.data(), x, r.data(), n, 1.0, 0.0);
  double bb = 0.0, rr = 0.0;
  for (int i = 0; i < n; i++) { bb += b[i] * b[i]; rr += (b[i] - r[i]) * (b[i] - r[i]); }
  *relres = bb > 0.0 ? std::sqrt(rr / bb) : 0.0;
  return 0;
}

// BCH:3808-3905
void step2(Case& c, int iter) {
  const int NN = c.NN;
  std::vector<double> dummy((size_t)3*NN);
  double o2 = 1.0000000000000000 / (c.dt * c.dt);
  for (size_t i = 0; i < (size_t)3*NN; i++) dummy[i] = o2 * c.UnpHalf[i];
  if (iter != 1)
    for (size_t i = 0; i < (size_t)3*NN; i++) dummy[i] = dummy[i] - c.MdOrigInv[i] * c.KAcc[i];
  if (c.f32_operands && c.mass_mode == 0)                              // the device's FP32 shadow of it
    for (size_t i = 0; i < (size_t)3*NN; i++) dummy[i] = (double)(float)dummy[i];
  const std::vector<double>* Gs[3] = {&c.G1, &c.G2, &c.G3};
  for (int d = 0; d < 3; d++)
    gtmv(c, Gs[d]->data(), dummy.data() + (size_t)d*NN, c.R2.data(), 1.0, d == 0 ? 0.0 : 1.0);
  if (c.zeroP > 0) c.R2[c.zeroP] = 0.0;                                  // BCH:4180-4184
  auto t0 = std::chrono::steady_clock::now();
  if (c.pcg_direct) {
    c.last_pcg_iters = 0;
    if (direct_solve(c, c.R2.data(), c.Pdot.data(), &c.last_pcg_relres) != 0)
      for (int i = 0; i < c.NNp; i++) c.Pdot[i] = NAN;
  } else {
    c.last_pcg_iters = pcg(c.Zrow, c.Zcol, c.Zval, c.NNp, c.R2.data(), c.Pdot.data(), c.pcg_rtol, c.pcg_maxit,
                           &c.last_pcg_relres, c.pcg_warm);
  }
  c.t_pcg_total += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  c.total_pcg_iters += c.last_pcg_iters;
  c.n_solves++;
  for (int i = 0; i < c.NNp; i++) c.Pnp1[i] = c.Pn[i] + c.dt * c.Pdot[i];
}

// BCH:3912-3974
void step3(Case& c, int iter) {
  const int NN = c.NN;
  const std::vector<double>* Gs[3] = {&c.G1, &c.G2, &c.G3};
  for (int d = 0; d < 3; d++)
    csrmv(c.Grow, c.Gcol, Gs[d]->data(), c.Pdot.data(), c.R3.data() + (size_t)d*NN, NN, -c.dt, 0.0);
  if (iter != 1)
    for (size_t i = 0; i < (size_t)3*NN; i++) c.R3[i] = c.R3[i] - c.dt * c.KAcc[i];
  for (int n : c.BCnodes) { c.R3[n] = 0.0; c.R3[n + NN] = 0.0; c.R3[n + 2*(size_t)NN] = 0.0; }   // BCH:4196-4206
  if (c.mass_mode == 1) mass_solve(c, c.R3.data(), c.Acc.data());
  else
    for (size_t i = 0; i < (size_t)3*NN; i++) c.Acc[i] = c.R3[i] * c.MdInv[i];
  for (size_t i = 0; i < (size_t)3*NN; i++) c.Unp1[i] = c.UnpHalf[i] + c.dt * c.Acc[i];
}

// One pass of the time loop body, BCH:2859-3118. Returns 1 if the
// steady-state criterion was met (the reference then stops).
int timeStep(Case& c) {
  auto t0 = std::chrono::steady_clock::now();
  const int NN = c.NN;
  c.timeN++;
  c.timeT += c.dt;
  c.UnpHalf_prev = c.Un;                                                 // BCH:2872-2876 (Unp1_prev not reset)
  std::fill(c.Acc_prev.begin(), c.Acc_prev.end(), 0.0);
  c.Pnp1_prev = c.Pn;
  int iter;
  for (iter = 1; iter <= c.maxIter; iter++) {
    step1(c, iter);
    step2(c, iter);
    step3(c, iter);
    double s1 = 0, s2 = 0;                                               // BCH:2936-2961
    for (size_t i = 0; i < (size_t)3*NN; i++) { s1 += c.Unp1[i]*c.Unp1[i]; double d = c.Unp1[i] - c.Unp1_prev[i]; s2 += d*d; }
    double nu = std::sqrt(s2) / std::sqrt(s1);
    s1 = 0; s2 = 0;
    for (int i = 0; i < c.NNp; i++) { s1 += c.Pnp1[i]*c.Pnp1[i]; double d = c.Pnp1[i] - c.Pnp1_prev[i]; s2 += d*d; }
    double np = std::sqrt(s2) / std::sqrt(s1);
    if (nu < c.tolerance && np < c.tolerance) break;
    c.UnpHalf_prev = c.UnpHalf;                                          // BCH:2964-2972
    c.Unp1_prev = c.Unp1;
    c.Acc_prev = c.Acc;
    c.Pnp1_prev = c.Pnp1;
    const std::vector<double> ak = k_operand(c, c.Acc_prev);
    for (int d = 0; d < 3; d++)                                          // BCH:2981-3035
      csrmv(c.Mrow, c.Mcol, c.Kval.data(), ak.data() + (size_t)d*NN, c.KAcc.data() + (size_t)d*NN, NN, 1.0, 0.0);
  }
  c.last_inner_iters = std::min(iter, c.maxIter);
  double oneOverdt = 1.0000000000000000 / c.dt;                         // BCH:3048-3062
  c.maxAcc = 0.0;
  for (size_t i = 0; i < (size_t)3*NN; i++) {
    double a = std::fabs((c.Unp1[i] - c.Un[i]) * oneOverdt);
    if (a > c.maxAcc) c.maxAcc = a;
  }
  int steady = c.maxAcc > c.converge ? 0 : 1;
  c.Un = c.Unp1;                                                         // BCH:3070-3077
  c.Pn = c.Pnp1;
  c.t_step_total += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return steady;
}

bool setup(Case& c) {
  if (!findElemsOfPresNodes(c)) return false;
  findElemNeighbors(c);
  if (!setupMeshColoring(c)) return false;
  setupNonCornerNodes(c);
  determineVelBCnodes(c);
  if (!findElemsOfVelNodes(c)) return false;
  findMonitorPoint(c);
  setupSparse(c);
  if (!setupGQ(c)) return false;
  calcShape(c);
  step0(c);
  calculateZ(c);
  initialize(c);
  return true;
}

std::vector<double>* dvec(Case& c, const char* name) {
  std::string n(name);
  if (n == "Un") return &c.Un;           if (n == "Unp1") return &c.Unp1;
  if (n == "Unp1_prev") return &c.Unp1_prev; if (n == "UnpHalf") return &c.UnpHalf;
  if (n == "UnpHalf_prev") return &c.UnpHalf_prev; if (n == "Acc") return &c.Acc;
  if (n == "Acc_prev") return &c.Acc_prev; if (n == "KAcc") return &c.KAcc;
  if (n == "Pn") return &c.Pn;           if (n == "Pnp1") return &c.Pnp1;
  if (n == "Pnp1_prev") return &c.Pnp1_prev; if (n == "Pdot") return &c.Pdot;
  if (n == "R1") return &c.R1;           if (n == "R2") return &c.R2;   if (n == "R3") return &c.R3;
  if (n == "Md") return &c.Md;           if (n == "MdOrig") return &c.MdOrig;
  if (n == "MdInv") return &c.MdInv;     if (n == "MdOrigInv") return &c.MdOrigInv;
  if (n == "K") return &c.Kval;          if (n == "M") return &c.Mval;
  if (n == "G1") return &c.G1;           if (n == "G2") return &c.G2;   if (n == "G3") return &c.G3;
  if (n == "Z") return &c.Zval;          if (n == "coord") return &c.coord;
  if (n == "Sv") return &c.Sv;           if (n == "Sp") return &c.Sp;
  if (n == "dSv") return &c.dSv;         if (n == "gqp") return &c.gqp; if (n == "gqw") return &c.gqw;
  return nullptr;
}

std::vector<int>* ivec(Case& c, const char* name) {
  std::string n(name);
  if (n == "LtoGnode") return &c.LtoGnode; if (n == "Mrow") return &c.Mrow; if (n == "Mcol") return &c.Mcol;
  if (n == "Grow") return &c.Grow;  if (n == "Gcol") return &c.Gcol;
  if (n == "Zrow") return &c.Zrow;  if (n == "Zcol") return &c.Zcol;
  if (n == "mapM") return &c.mapM;  if (n == "mapG") return &c.mapG;
  if (n == "BCnodes") return &c.BCnodes; if (n == "BCwhich") return &c.BCwhich;
  if (n == "meshColors") return &c.meshColors; if (n == "elementsOfColor") return &c.elementsOfColor;
  if (n == "neigh") return &c.neigh; if (n == "nNeigh") return &c.nNeigh;
  return nullptr;
}

}  // namespace

// ============================================================ C API (ctypes)
extern "C" {

void* orc_new(void) { return new Case(); }
void orc_free(void* h) { delete (Case*)h; }
const char* orc_error(void* h) { return ((Case*)h)->err.c_str(); }

int orc_read_inp(void* h, const char* path) { return read_inp(*(Case*)h, path) ? 0 : 1; }

// Mesh from arrays: conn8 is 0-based NE*8; faces is nFaces*3 0-based (elem, face, bc).
int orc_set_mesh(void* h, int NE, int NCN, const double* coord, const int* conn8, int nBC, const double* BCstr,
                 int nFaces, const int* faces, int zeroP, const double* mon) {
  Case& c = *(Case*)h;
  c.NE = NE; c.NCN = NCN; c.NNp = NCN;
  c.coord.assign(coord, coord + (size_t)NCN * 3);
  c.LtoGnode.assign((size_t)NE * 27, -1);
  for (int e = 0; e < NE; e++) for (int i = 0; i < 8; i++) c.LtoGnode[(size_t)e*27 + i] = conn8[(size_t)e*8 + i];
  c.nBC = nBC; c.BCstr.assign(BCstr, BCstr + (size_t)nBC * 3);
  c.velFaces.assign(faces, faces + (size_t)nFaces * 3);
  c.zeroP = zeroP;
  for (int d = 0; d < 3; d++) c.mon[d] = mon[d];
  return 0;
}

// params: NGP, density, viscosity, dt, maxIter, tolerance, converge, pcg_rtol, pcg_maxit
int orc_set_params(void* h, int NGP, double density, double viscosity, double dt, int maxIter, double tolerance,
                   double converge, double pcg_rtol, int pcg_maxit) {
  Case& c = *(Case*)h;
  c.NGP = NGP; c.density = density; c.viscosity = viscosity; c.dt = dt; c.maxIter = maxIter;
  c.tolerance = tolerance; c.converge = converge; c.pcg_rtol = pcg_rtol; c.pcg_maxit = pcg_maxit;
  return 0;
}

int orc_get_header(void* h, double* out) {  // NGP, density, viscosity, dt, t_ini, t_final, maxIter, tol, converge
  Case& c = *(Case*)h;
  double v[9] = {(double)c.NGP, c.density, c.viscosity, c.dt, c.t_ini, c.t_final, (double)c.maxIter, c.tolerance, c.converge};
  std::memcpy(out, v, sizeof v);
  return 0;
}

int orc_setup(void* h) {
  Case& c = *(Case*)h;
  omp_set_dynamic(0);
  return setup(c) ? 0 : 1;
}

// sizes: NE, NCN, NN, NNp, nnzM, nnzG, nnzZ, nBCnodes, nActiveColors, monPoint, zeroP, nVelFaces, nBC
int orc_sizes(void* h, long long* out) {
  Case& c = *(Case*)h;
  long long v[13] = {c.NE, c.NCN, c.NN, c.NNp, (long long)c.Mcol.size(), (long long)c.Gcol.size(), (long long)c.Zcol.size(),
                     (long long)c.BCnodes.size(), c.nActiveColors, c.monPoint, c.zeroP, (long long)c.velFaces.size() / 3, c.nBC};
  std::memcpy(out, v, sizeof v);
  return 0;
}

long long orc_len(void* h, const char* name, int is_int) {
  Case& c = *(Case*)h;
  if (is_int) { auto* v = ivec(c, name); return v ? (long long)v->size() : -1; }
  auto* v = dvec(c, name); return v ? (long long)v->size() : -1;
}
int orc_get_d(void* h, const char* name, double* dst) {
  auto* v = dvec(*(Case*)h, name); if (!v) return 1;
  std::memcpy(dst, v->data(), v->size() * sizeof(double)); return 0;
}
int orc_set_d(void* h, const char* name, const double* src) {
  auto* v = dvec(*(Case*)h, name); if (!v) return 1;
  std::memcpy(v->data(), src, v->size() * sizeof(double));
  if (v == &((Case*)h)->Zval) ((Case*)h)->chol_w = -1;   // the kept factor is of the old Z
  return 0;
}
int orc_get_i(void* h, const char* name, int* dst) {
  auto* v = ivec(*(Case*)h, name); if (!v) return 1;
  std::memcpy(dst, v->data(), v->size() * sizeof(int)); return 0;
}

// Run nsteps time steps (or until steady state). stats per step (6 doubles):
// inner iterations, total PCG iterations in the step, maxAcc, u/v/w/p monitor
// are written as [iters, pcg, maxAcc, um, vm, wm, pm, wall_s] -> 8 doubles.
int orc_run(void* h, int nsteps, double* stats) {
  Case& c = *(Case*)h;
  for (int s = 0; s < nsteps; s++) {
    long before = c.total_pcg_iters;
    double tb = c.t_step_total;
    int steady = timeStep(c);
    if (stats) {
      double* o = stats + 8 * s;
      o[0] = c.last_inner_iters; o[1] = (double)(c.total_pcg_iters - before); o[2] = c.maxAcc;
      o[3] = c.Un[c.monPoint]; o[4] = c.Un[c.monPoint + c.NN]; o[5] = c.Un[c.monPoint + 2*(size_t)c.NN];
      o[6] = c.Pn[c.monPoint]; o[7] = c.t_step_total - tb;
    }
    if (steady) return s + 1;
  }
  return nsteps;
}

int orc_step1(void* h, int iter) { step1(*(Case*)h, iter); return 0; }

// pressure PCG initial guess: 0 zero (BCH:4002-4004), 1 the previous Pdot (CU:425-442)
int orc_set_pcg_x0(void* h, int policy) {
  Case& c = *(Case*)h;
  if (policy < 0 || policy > 1) { c.err = "orc_set_pcg_x0: bad policy"; return 1; }
  c.pcg_warm = policy == 1;
  return 0;
}

// pressure solve: 0 Jacobi-PCG (the reference), 1 Cholesky of Z (Case::pcg_direct)
int orc_set_direct(void* h, int on) {
  Case& c = *(Case*)h;
  if (on < 0 || on > 1) { c.err = "orc_set_direct: 0 or 1"; return 1; }
  c.pcg_direct = on == 1;
  return 0;
}
// x = Z^-1 b by the kept (or a new) Cholesky factor; relres = |b - Z x| / |b|
int orc_direct_solve(void* h, const double* b, double* x, double* relres) {
  return direct_solve(*(Case*)h, b, x, relres);
}
// half-bandwidth of the kept factor (-1 when none)
int orc_chol_bandwidth(void* h) { return ((Case*)h)->chol_w; }

// Symbolic Cholesky of the symmetric pattern (rowptr, col; both triangles,
// diagonal included) under the order perm (perm[new] = old): the column
// counts of L (diagonal included), summed into *nnzL, with sum_k c_k^2 in
// *flops (the multiply-add count of the numeric factor, up to lower-order
// terms) and the elimination tree's height. The elimination tree is built
// with path compression (ancestor links); row i of L is the union of the
// tree paths from the row's entries j < i up to i, so walking each path
// until a node already marked for row i counts every entry of L once.
// (The prototype orders Z with symamd before chol, ML:1880-1881; this is
// the fill any order gives, for comparing orders.)
int orc_symbolic_cholesky(int n, const int* rowptr, const int* col, const int* perm, long long* nnzL, double* flops,
                          int* height) {
  if (n <= 0 || !rowptr || !col || !perm) return 1;
  std::vector<int> iperm(n, -1), parent(n, -1), anc(n, -1), mark(n, -1);
  for (int i = 0; i < n; i++) {
    if (perm[i] < 0 || perm[i] >= n || iperm[perm[i]] >= 0) return 2;
    iperm[perm[i]] = i;
  }
  for (int i = 0; i < n; i++) {
    const int o = perm[i];
    for (int t = rowptr[o]; t < rowptr[o + 1]; t++) {
      int k = iperm[col[t]];
      while (k >= 0 && k < i) {
        const int nx = anc[k];
        anc[k] = i;
        if (nx < 0) { parent[k] = i; break; }
        k = nx;
      }
    }
  }
  std::vector<long long> cc(n, 1);
  for (int i = 0; i < n; i++) {
    mark[i] = i;
    const int o = perm[i];
    for (int t = rowptr[o]; t < rowptr[o + 1]; t++) {
      int k = iperm[col[t]];
      if (k >= i) continue;
      while (mark[k] != i) {
        mark[k] = i;
        cc[k]++;
        k = parent[k];
      }
    }
  }
  long long s = 0;
  double f = 0.0;
  for (int k = 0; k < n; k++) { s += cc[k]; f += (double)cc[k] * (double)cc[k]; }
  std::vector<int> depth(n, 0);
  int h = 0;
  for (int k = n - 1; k >= 0; k--) {
    depth[k] = parent[k] < 0 ? 1 : depth[parent[k]] + 1;
    h = std::max(h, depth[k]);
  }
  if (nnzL) *nnzL = s;
  if (flops) *flops = f;
  if (height) *height = h;
  return 0;
}

// step1's K operands and step2's G^T operand in FP64 (0) or rounded to FP32
// (1); see Case::f32_operands
int orc_set_f32_operands(void* h, int f32) {
  Case& c = *(Case*)h;
  if (f32 < 0 || f32 > 1) { c.err = "orc_set_f32_operands: 0 or 1"; return 1; }
  c.f32_operands = f32 == 1;
  return 0;
}

// velocity mass mode (0 lumped, 1 consistent with Jacobi-PCG per component)
int orc_set_mass(void* h, int mode, double rtol, int maxit) {
  Case& c = *(Case*)h;
  if (mode < 0 || mode > 1 || !(rtol >= 0) || maxit < 1) { c.err = "orc_set_mass: bad arguments"; return 1; }
  c.mass_mode = mode; c.mass_rtol = rtol; c.mass_maxit = maxit;
  return 0;
}
// last mass solve: largest iteration count and relative residual over the components; total iterations
int orc_mass_stats(void* h, double* out) {
  Case& c = *(Case*)h;
  out[0] = c.last_mass_iters; out[1] = c.last_mass_relres; out[2] = (double)c.total_mass_iters;
  return 0;
}
int orc_step2(void* h, int iter) { step2(*(Case*)h, iter); return 0; }
int orc_step3(void* h, int iter) { step3(*(Case*)h, iter); return 0; }
int orc_kacc(void* h) {
  Case& c = *(Case*)h;
  for (int d = 0; d < 3; d++)
    csrmv(c.Mrow, c.Mcol, c.Kval.data(), c.Acc.data() + (size_t)d*c.NN, c.KAcc.data() + (size_t)d*c.NN, c.NN, 1.0, 0.0);
  return 0;
}
int orc_convection(void* h, const double* Un, const double* Uprev, double* R1) {
  convection(*(Case*)h, Un, Uprev, R1); return 0;
}
// Solve Z x = b with the case's Z.
int orc_pcg_z(void* h, const double* b, double* x, double rtol, int maxit, double* relres) {
  Case& c = *(Case*)h;
  return pcg(c.Zrow, c.Zcol, c.Zval, c.NNp, b, x, rtol, maxit, relres);
}
// Generic CSR Jacobi-PCG.
int orc_pcg_csr(int n, const int* rp, const int* ci, const double* v, const double* b, double* x, double rtol, int maxit,
                double* relres) {
  std::vector<int> r(rp, rp + n + 1), cc(ci, ci + rp[n]);
  std::vector<double> vv(v, v + rp[n]);
  return pcg(r, cc, vv, n, b, x, rtol, maxit, relres);
}
// Which: 0 = K (NN x NN), 1..3 = G_d (NN x NNp), 4 = Z. y = alpha*A x + beta*y
int orc_spmv(void* h, int which, const double* x, double* y, double alpha, double beta) {
  Case& c = *(Case*)h;
  if (which == 0) csrmv(c.Mrow, c.Mcol, c.Kval.data(), x, y, c.NN, alpha, beta);
  else if (which >= 1 && which <= 3) {
    const std::vector<double>* Gs[3] = {&c.G1, &c.G2, &c.G3};
    csrmv(c.Grow, c.Gcol, Gs[which-1]->data(), x, y, c.NN, alpha, beta);
  } else if (which == 4) csrmv(c.Zrow, c.Zcol, c.Zval.data(), x, y, c.NNp, alpha, beta);
  else if (which >= 5 && which <= 7) {
    const std::vector<double>* Gs[3] = {&c.G1, &c.G2, &c.G3};
    gtmv(c, Gs[which-5]->data(), x, y, alpha, beta);
  } else return 1;
  return 0;
}
// Timings: total step wall time, total PCG wall time, total PCG iterations, number of solves
int orc_timings(void* h, double* out) {
  Case& c = *(Case*)h;
  out[0] = c.t_step_total; out[1] = c.t_pcg_total; out[2] = (double)c.total_pcg_iters; out[3] = c.n_solves;
  return 0;
}
int orc_set_threads(int n) { omp_set_num_threads(n); return 0; }

}  // extern "C"
