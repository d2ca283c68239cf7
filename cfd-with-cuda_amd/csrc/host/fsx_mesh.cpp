// fsx_mesh.cpp -- mesh input for the fractional-step core: the reference's
// .inp layout (reader BCH:528-725) and synthetic generators for the lid-driven
// cavity (layout of oldFiles/meshGenerators&Converters/
// HexaMeshGeneratorInACube_GeneratesCornerNodes.m, "GEN") and the bent square
// duct (channel generator HexaMeshGeneratorInAChannel_GeneratesCornerNodes.m,
// "GENCH", bent and optionally renumbered).
#include "fsx_host_internal.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <numeric>
#include <random>
#include <sstream>
#include <string>

namespace fsx {

// GEN:100-131: uniform or sinh-clustered 1-D node positions on [0, L]
static std::vector<double> axis_nodes(int n_el, double L, double cluster) {
  const int nc = n_el + 1;
  std::vector<double> d(nc);
  if (cluster == 0.0) {
    for (int i = 0; i < nc; i++) d[i] = (i * L) / (nc - 1);
    return d;
  }
  const double mx = std::sinh(cluster);
  const int half = (nc + 1) / 2;
  for (int i = 1; i <= half; i++) {
    double xx = (1.0 * (i - 1)) / (half - 1);
    d[i - 1] = L / 2.0 / mx * std::sinh(cluster * xx);
  }
  for (int i = half; i <= nc; i++) d[i - 1] = L - d[nc + 1 - i - 1];
  return d;
}

}  // namespace fsx

using namespace fsx;

extern "C" fsxh_mesh* fsxh_mesh_cavity(int N, double cluster, double lid_u) {
  // the mirrored sinh clustering of GEN:100-131 needs an even element count
  // ("Select an even number for proper clustering", GEN:41)
  if (N < 1 || (cluster != 0.0 && N % 2 != 0)) return nullptr;
  auto* m = new fsxh_mesh();
  const int nx = N, ny = N, nz = N, ncx = N + 1, ncy = N + 1, ncz = N + 1;
  auto dx = axis_nodes(nx, 1.0, cluster), dy = axis_nodes(ny, 1.0, cluster), dz = axis_nodes(nz, 1.0, cluster);
  m->NCN = ncx * ncy * ncz;
  m->NE = nx * ny * nz;
  m->coord.resize((size_t)m->NCN * 3);
  for (int k = 0; k < ncz; k++)                                          // GEN:134-148, x fastest
    for (int j = 0; j < ncy; j++)
      for (int i = 0; i < ncx; i++) {
        size_t n = ((size_t)k * ncy + j) * ncx + i;
        m->coord[3*n] = dx[i]; m->coord[3*n+1] = dy[j]; m->coord[3*n+2] = dz[k];
      }
  m->conn.resize((size_t)m->NE * 8);
  const int lay = ncx * ncy;
  for (int k = 0; k < nz; k++)                                           // GEN:152-169
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        size_t e = ((size_t)k * ny + j) * nx + i;
        int n1 = k * lay + j * ncx + i, n2 = n1 + 1, n3 = n1 + nx + 2, n4 = n3 - 1;
        int c[8] = {n1, n2, n3, n4, n1 + lay, n2 + lay, n3 + lay, n4 + lay};
        std::memcpy(&m->conn[e*8], c, sizeof c);
      }
  auto el = [&](int i, int j, int k) { return (k * ny + j) * nx + i; };
  auto face = [&](int e, int f, int b) { m->faces.push_back(e); m->faces.push_back(f); m->faces.push_back(b); };
  for (int j = 0; j < ny; j++) for (int i = 0; i < nx; i++) face(el(i, j, 0), 0, 0);        // bottom
  for (int k = 0; k < nz; k++) for (int j = 0; j < ny; j++) face(el(0, j, k), 4, 0);        // left
  for (int k = 0; k < nz; k++) for (int j = 0; j < ny; j++) face(el(nx - 1, j, k), 2, 0);   // right
  for (int k = 0; k < nz; k++) for (int i = 0; i < nx; i++) face(el(i, 0, k), 1, 0);        // front
  for (int k = 0; k < nz; k++) for (int i = 0; i < nx; i++) face(el(i, ny - 1, k), 3, 0);   // back
  for (int j = 0; j < ny; j++) for (int i = 0; i < nx; i++) face(el(i, j, nz - 1), 5, 1);   // lid, last (Q7)
  m->nBC = 2;
  m->BCstr = {0.0, 0.0, 0.0, lid_u, 0.0, 0.0};
  m->zeroP = (ny / 2) * ncx + (ncx + 1) / 2 - 1;                        // GEN:294, 0-based
  m->mon[0] = m->mon[1] = m->mon[2] = 0.5;
  m->title = "3D Lid-driven cavity problem";
  return m;
}

extern "C" fsxh_mesh* fsxh_mesh_duct(int n, int n_in, int n_arc, int n_out, double L_in, double R, double L_out,
                                     double peak, uint64_t seed, int permute) {
  if (n < 1 || n_in < 0 || n_arc < 1 || n_out < 0 || R <= 0.5) return nullptr;
  auto* m = new fsxh_mesh();
  const int na = n_in + n_arc + n_out;          // axial elements (x-like, fastest)
  const int nca = na + 1, ncs = n + 1, nct = n + 1;
  m->NCN = nca * ncs * nct;
  m->NE = na * n * n;
  const double pi = std::acos(-1.0);
  // centreline frame: inlet along +x ending at the origin, quarter bend about
  // (0, R) ending at (R, R), outlet along +y; s offset in [-0.5, 0.5] along the
  // in-plane normal, t = z in [0, 1]
  auto position = [&](int a, double s_off, double t, double* p) {
    if (a <= n_in) {
      double x = -L_in + L_in * a / (double)n_in;
      if (n_in == 0) x = 0.0;
      p[0] = x; p[1] = s_off; p[2] = t;
    } else if (a <= n_in + n_arc) {
      double th = (pi / 2) * (a - n_in) / (double)n_arc;
      p[0] = (R - s_off) * std::sin(th); p[1] = R - (R - s_off) * std::cos(th); p[2] = t;
    } else {
      double yb = L_out * (a - n_in - n_arc) / (double)n_out;
      p[0] = R - s_off; p[1] = R + yb; p[2] = t;
    }
  };
  std::vector<double> xyz((size_t)m->NCN * 3);
  for (int k = 0; k < nct; k++)
    for (int j = 0; j < ncs; j++)
      for (int i = 0; i < nca; i++) {
        size_t id = ((size_t)k * ncs + j) * nca + i;
        position(i, -0.5 + j / (double)n, k / (double)n, &xyz[3*id]);
      }
  std::vector<int> conn((size_t)m->NE * 8);
  const int lay = nca * ncs;
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < na; i++) {
        size_t e = ((size_t)k * n + j) * na + i;
        int n1 = k * lay + j * nca + i, n2 = n1 + 1, n3 = n1 + na + 2, n4 = n3 - 1;
        int c[8] = {n1, n2, n3, n4, n1 + lay, n2 + lay, n3 + lay, n4 + lay};
        std::memcpy(&conn[e*8], c, sizeof c);
      }
  std::vector<int> faces;
  auto el = [&](int i, int j, int k) { return (k * n + j) * na + i; };
  auto face = [&](int e, int f, int b) { faces.push_back(e); faces.push_back(f); faces.push_back(b); };
  // GENCH:218-262 order: bottom, front, back, top walls (BC 1), inlet last (BC 2)
  for (int j = 0; j < n; j++) for (int i = 0; i < na; i++) face(el(i, j, 0), 0, 0);
  for (int k = 0; k < n; k++) for (int i = 0; i < na; i++) face(el(i, 0, k), 1, 0);
  for (int k = 0; k < n; k++) for (int i = 0; i < na; i++) face(el(i, n - 1, k), 3, 0);
  for (int j = 0; j < n; j++) for (int i = 0; i < na; i++) face(el(i, j, n - 1), 5, 0);
  for (int k = 0; k < n; k++) for (int j = 0; j < n; j++) face(el(0, j, k), 4, 1);
  int zeroP = (n / 2) * nca * ncs + (n / 2 + 1) * nca - 1;               // GENCH:287, 0-based
  // optional renumbering: node ids and element order
  std::vector<int> nperm(m->NCN), eperm(m->NE);
  std::iota(nperm.begin(), nperm.end(), 0);
  std::iota(eperm.begin(), eperm.end(), 0);
  if (permute) {
    std::mt19937_64 rng(seed);
    for (int i = m->NCN - 1; i > 0; i--) { std::uniform_int_distribution<int> d(0, i); std::swap(nperm[i], nperm[d(rng)]); }
    for (int i = m->NE - 1; i > 0; i--) { std::uniform_int_distribution<int> d(0, i); std::swap(eperm[i], eperm[d(rng)]); }
  }
  // nperm[old] = new node id, eperm[new element slot] = old element
  std::vector<int> einv(m->NE);
  for (int e = 0; e < m->NE; e++) einv[eperm[e]] = e;
  m->coord.resize(xyz.size());
  for (int i = 0; i < m->NCN; i++) for (int d = 0; d < 3; d++) m->coord[3*(size_t)nperm[i] + d] = xyz[3*(size_t)i + d];
  m->conn.resize(conn.size());
  for (int e = 0; e < m->NE; e++) for (int c = 0; c < 8; c++) m->conn[(size_t)e*8 + c] = nperm[conn[(size_t)eperm[e]*8 + c]];
  for (size_t f = 0; f < faces.size(); f += 3) faces[f] = einv[faces[f]];
  m->faces = faces;
  m->zeroP = nperm[zeroP];
  m->nBC = 2;
  m->BCstr = {0.0, 0.0, 0.0, peak, 0.0, 0.0};
  m->has_profile = 1;
  m->profile_bc = 1; m->profile_peak = peak; m->profile_y0 = -0.5; m->profile_z0 = 0.0;
  double p[3];
  position(n_in + n_arc / 2, 0.0, 0.5, p);
  m->mon[0] = p[0]; m->mon[1] = p[1]; m->mon[2] = p[2];
  m->title = "3D bent square duct";
  m->viscosity = 1.0 / 200.0;
  return m;
}

extern "C" void fsxh_mesh_free(fsxh_mesh* m) { delete m; }

static double after_colon(const std::string& s, bool* ok) {
  size_t p = s.find(':');
  if (p == std::string::npos) { *ok = false; return 0; }
  char* end = nullptr;
  double v = std::strtod(s.c_str() + p + 1, &end);
  if (end == s.c_str() + p + 1) *ok = false;
  return v;
}

extern "C" int fsxh_mesh_read_inp(const char* path, fsxh_mesh** out, char* err, int errlen) {
  auto fail = [&](const std::string& msg) {
    if (err && errlen > 0) std::snprintf(err, errlen, "%s", msg.c_str());
    return (int)FSX_EINVAL;
  };
  std::ifstream f(path);
  if (!f) return fail(std::string("cannot open ") + path);
  auto m = std::make_unique<fsxh_mesh>();
  std::string s;
  std::getline(f, m->title);
  std::getline(f, s);
  double h[18];
  for (int i = 0; i < 18; i++) {                                         // BCH:547-564 (header order)
    bool ok = true;
    if (!std::getline(f, s)) return fail("truncated header");
    h[i] = after_colon(s, &ok);
    if (!ok) return fail("bad header line: " + s);
  }
  if ((int)h[0] != 1 || (int)h[3] != 27 || (int)h[4] != 8) return fail("only hex27 velocity / hex8 pressure supported");
  m->NE = (int)h[1]; m->NCN = (int)h[2]; m->NGP = (int)h[5];
  m->dt = h[7]; m->t_ini = h[8]; m->t_final = h[9]; m->maxIter = (int)h[10];
  m->tolerance = h[11]; m->converge = h[12]; m->density = h[14]; m->viscosity = h[15];
  if (m->NE <= 0 || m->NCN <= 0) return fail("bad NE/NCN");
  std::getline(f, s); std::getline(f, s);
  m->coord.resize((size_t)m->NCN * 3);
  for (int i = 0; i < m->NCN; i++) {
    if (!std::getline(f, s)) return fail("truncated coordinates");
    std::istringstream ls(s); int id;
    if (!(ls >> id >> m->coord[3*(size_t)i] >> m->coord[3*(size_t)i+1] >> m->coord[3*(size_t)i+2])) return fail("bad coordinate line");
  }
  std::getline(f, s); std::getline(f, s);
  m->conn.resize((size_t)m->NE * 8);
  for (int e = 0; e < m->NE; e++) {
    if (!std::getline(f, s)) return fail("truncated connectivity");
    std::istringstream ls(s); int id; ls >> id;
    for (int c = 0; c < 8; c++) {
      int v;
      if (!(ls >> v)) return fail("bad connectivity line");
      if (v < 1 || v > m->NCN) return fail("corner index out of range");
      m->conn[(size_t)e*8 + c] = v - 1;
    }
  }
  std::getline(f, s); std::getline(f, s);
  bool ok = true;
  std::getline(f, s); m->nBC = (int)after_colon(s, &ok);
  if (!ok || m->nBC < 0) return fail("bad nBC");
  m->BCstr.assign((size_t)m->nBC * 3, 0.0);
  for (int b = 0; b < m->nBC; b++) {                                     // BCH:641-653
    std::getline(f, s);
    size_t p = s.find(':');
    if (p == std::string::npos) return fail("bad BC line");
    std::string rest = s.substr(p + 1);
    for (char& ch : rest) if (ch == ':') ch = ' ';
    std::istringstream ls(rest); double type;
    if (!(ls >> type >> m->BCstr[3*b] >> m->BCstr[3*b+1] >> m->BCstr[3*b+2])) return fail("bad BC values");
  }
  std::getline(f, s);
  std::getline(f, s); int nVel = (int)after_colon(s, &ok);
  std::getline(f, s); int nOut = (int)after_colon(s, &ok);
  if (!ok || nVel < 0 || nOut < 0) return fail("bad face counts");
  std::getline(f, s); std::getline(f, s);
  m->faces.resize((size_t)nVel * 3);
  for (int i = 0; i < nVel; i++) {
    if (!std::getline(f, s)) return fail("truncated velocity BC faces");
    std::istringstream ls(s); int a, b, c;
    if (!(ls >> a >> b >> c)) return fail("bad velocity BC line");
    if (a < 1 || a > m->NE || b < 1 || b > 6 || c < 1 || c > m->nBC) return fail("velocity BC index out of range");
    m->faces[3*(size_t)i] = a - 1; m->faces[3*(size_t)i+1] = b - 1; m->faces[3*(size_t)i+2] = c - 1;
  }
  std::getline(f, s); std::getline(f, s);
  for (int i = 0; i < nOut; i++) std::getline(f, s);
  std::getline(f, s); std::getline(f, s);
  if (!std::getline(f, s)) return fail("missing zero-pressure node");
  m->zeroP = std::atoi(s.c_str()) - 1;
  std::getline(f, s); std::getline(f, s);
  if (!std::getline(f, s)) return fail("missing monitor point");
  { std::istringstream ls(s); ls >> m->mon[0] >> m->mon[1] >> m->mon[2]; }
  *out = m.release();
  return FSX_OK;
}

extern "C" int fsxh_mesh_write_inp(const fsxh_mesh* m, const char* path) {
  FILE* f = std::fopen(path, "w");
  if (!f) return FSX_EINVAL;
  std::fprintf(f, "%s\n================================================\n", m->title.c_str());
  std::fprintf(f, "eType    : 1 \nNE       : %d \nNCN      : %d \nNENv     : 27 \nNENp     : 8 \nNGP      : %d \n",
               m->NE, m->NCN, m->NGP);
  std::fprintf(f, "alpha    : 1.000000\ndt       : %.10g\nt_ini    : %.10g \nt_final  : %.10g \nmaxIter  : %d \n",
               m->dt, m->t_ini, m->t_final, m->maxIter);
  std::fprintf(f, "tolerance: %.10g\nconverge : %.10g \nisRestart: 0\ndensity  : %.10g \nviscosity: %.10g \n",
               m->tolerance, m->converge, m->density, m->viscosity);
  std::fprintf(f, "fx       : 0.0 \nfy       : 0.0 \n================================================\n");
  std::fprintf(f, "Corner Node No         x                y                z\n");
  for (int i = 0; i < m->NCN; i++)
    std::fprintf(f, "%9d %.17g %.17g %.17g\n", i + 1, m->coord[3*(size_t)i], m->coord[3*(size_t)i+1], m->coord[3*(size_t)i+2]);
  std::fprintf(f, "================================================\n");
  std::fprintf(f, "Elem No   corner1  corner2  corner3  corner4  corner5  corner6  corner7  corner8\n");
  for (int e = 0; e < m->NE; e++) {
    std::fprintf(f, "%6d ", e + 1);
    for (int c = 0; c < 8; c++) std::fprintf(f, " %8d", m->conn[(size_t)e*8 + c] + 1);
    std::fprintf(f, " \n");
  }
  std::fprintf(f, "================================================\nBCs (Number of specified BCs, their types and strings) \n");
  std::fprintf(f, "nBC       : %d \n", m->nBC);
  for (int b = 0; b < m->nBC; b++)
    std::fprintf(f, "BC %d      : 1  %.17g : %.17g : %.17g\n", b + 1, m->BCstr[3*b], m->BCstr[3*b+1], m->BCstr[3*b+2]);
  std::fprintf(f, "================================================\nnVelFaces : %zu \nnOutFaces : 0 \n", m->faces.size() / 3);
  std::fprintf(f, "================================================\nVelocity BC (Elem# Face# BC#)\n");
  for (size_t i = 0; i < m->faces.size(); i += 3)
    std::fprintf(f, "%5d    %d    %d\n", m->faces[i] + 1, m->faces[i+1] + 1, m->faces[i+2] + 1);
  std::fprintf(f, "================================================\nOutflow BC (Elem# Face# BC#)\n");
  std::fprintf(f, "================================================\nNode number where pressure is taken to be zero\n%d\n",
               m->zeroP + 1);
  std::fprintf(f, "================================================\nMonitor point coordinates\n%.17g  %.17g  %.17g\n",
               m->mon[0], m->mon[1], m->mon[2]);
  std::fclose(f);
  return FSX_OK;
}

extern "C" int fsxh_mesh_info(const fsxh_mesh* m, long long* out) {
  long long v[8] = {m->NE, m->NCN, (long long)m->faces.size() / 3, m->nBC, m->zeroP, m->NGP, m->maxIter, m->has_profile};
  std::memcpy(out, v, sizeof v);
  return FSX_OK;
}

extern "C" int fsxh_mesh_params(const fsxh_mesh* m, double* out) {
  double v[10] = {m->dt, m->t_ini, m->t_final, m->tolerance, m->converge, m->density, m->viscosity,
                  m->mon[0], m->mon[1], m->mon[2]};
  std::memcpy(out, v, 10 * sizeof(double));
  return FSX_OK;
}

extern "C" int fsxh_mesh_set_params(fsxh_mesh* m, int NGP, double dt, int maxIter, double tolerance, double converge,
                                    double density, double viscosity) {
  if (NGP != 1 && NGP != 8 && NGP != 27) return FSX_EINVAL;
  if (!(dt > 0) || maxIter < 1 || !(density > 0) || !(viscosity >= 0)) return FSX_EINVAL;
  m->NGP = NGP; m->dt = dt; m->maxIter = maxIter; m->tolerance = tolerance; m->converge = converge;
  m->density = density; m->viscosity = viscosity;
  return FSX_OK;
}

extern "C" int fsxh_mesh_get(const fsxh_mesh* m, const char* name, void* dst) {
  std::string n(name);
  if (n == "coord") std::memcpy(dst, m->coord.data(), m->coord.size() * sizeof(double));
  else if (n == "conn") std::memcpy(dst, m->conn.data(), m->conn.size() * sizeof(int));
  else if (n == "faces") std::memcpy(dst, m->faces.data(), m->faces.size() * sizeof(int));
  else if (n == "BCstr") std::memcpy(dst, m->BCstr.data(), m->BCstr.size() * sizeof(double));
  else return FSX_EINVAL;
  return FSX_OK;
}

extern "C" int fsxh_mesh_profile(const fsxh_mesh* m, double* out) {
  out[0] = m->has_profile ? m->profile_bc : -1;
  out[1] = m->profile_peak; out[2] = m->profile_y0; out[3] = m->profile_z0;
  return FSX_OK;
}
