// Internal host-side structures of the fractional-step core.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "../../../include/fsx_host.h"

// Mesh input: the content of a reference .inp file (BCH:528-725).
struct fsxh_mesh {
  std::string title;
  int NE = 0, NCN = 0, NGP = 8, maxIter = 3;
  double dt = 0.01, t_ini = 0.0, t_final = 1.0, tolerance = 1e-3, converge = 1e-6;
  double density = 1.0, viscosity = 0.01;
  std::vector<double> coord;   // NCN*3
  std::vector<int> conn;       // NE*8, 0-based
  std::vector<int> faces;      // nVelFaces*3, 0-based (elem, face, bc)
  int nBC = 0;
  std::vector<double> BCstr;   // nBC*3
  int zeroP = -1;              // 0-based
  double mon[3] = {0.5, 0.5, 0.5};
  // optional parabolic inlet (the commented profile of BCH:4086-4102, with
  // peak 16 s(1-s) t(1-t) scaling) for BC id profile_bc
  int has_profile = 0, profile_bc = -1;
  double profile_peak = 1.0, profile_y0 = 0.0, profile_z0 = 0.0;
};

// Topology: numbering, BC nodes, CSR patterns, colouring.
struct fsxh_topo {
  int NE = 0, NCN = 0, NN = 0, NNp = 0, monPoint = 0, zeroP = -1;
  std::vector<int> LtoGnode;   // NE*27
  std::vector<double> coord;   // NN*3
  std::vector<int> BCnodes, BCwhich;
  std::vector<double> BCvalues;  // nBC*3 per BC node
  std::vector<int> Mrow, Mcol, Grow, Gcol, Zrow, Zcol;
  std::vector<int> color_ptr, color_order;
};

struct fsxd_solver {
  fsxh_topo* topo = nullptr;
  fsx_ctx* ctx = nullptr;
  fsxd_options opt{};
  double dt = 0.0, t = 0.0, t0 = 0.0;
  int timeN = 0;
  int steady = 0;
  double t_topo = 0, t_upload = 0, t_assemble = 0;
  std::string err;
};

namespace fsx {
// Local node order of the hex27 element (BCH:2312-2343): 0-7 corners,
// 8-19 edges, 20-25 faces, 26 centre.
extern const int kEdge[12][2];
extern const int kFace[6][4];
extern const int kFaceMid[6][5];
}  // namespace fsx
