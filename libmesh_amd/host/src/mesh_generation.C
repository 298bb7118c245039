// mesh_generation.C -- MeshTools::Generation::build_cube for the element types
// of the assembly path, the first-touch renumbering of prepare_for_use, and the
// seeded synthetic inputs (interior perturbation, log-normal coefficients).
#include <algorithm>
#include <cmath>
#include <map>
#include <random>

#include "libmesh/replicated_mesh.h"

namespace libMesh {
namespace {

// Lagrange node -> tensor index tables (fe_lagrange_shape_3D.C:277-279, 348-350);
// index 0 is the -1 end, 1 the +1 end, 2 the midpoint.  The lattice offset of a
// HEX27 node inside its element is {0, 2, 1}[index] half-cells.
const int kHex8Idx[3][8] = {{0, 1, 1, 0, 0, 1, 1, 0}, {0, 0, 1, 1, 0, 0, 1, 1}, {0, 0, 0, 0, 1, 1, 1, 1}};
const int kHex27Idx[3][27] = {
    {0, 1, 1, 0, 0, 1, 1, 0, 2, 1, 2, 0, 0, 1, 1, 0, 2, 1, 2, 0, 2, 2, 1, 2, 0, 2, 2},
    {0, 0, 1, 1, 0, 0, 1, 1, 0, 2, 1, 2, 0, 0, 1, 1, 0, 2, 1, 2, 2, 0, 2, 1, 2, 2, 2},
    {0, 0, 0, 0, 1, 1, 1, 1, 0, 0, 0, 0, 2, 2, 2, 2, 1, 1, 1, 1, 0, 2, 2, 2, 2, 1, 2}};
const int kHalfOffset[3] = {0, 2, 1};
// Vertex cycle of each hex side in side_nodes_map order (cell_hex27.C:37-45);
// the side's centre node is 20 + side.
const int kSideCycle[6][4] = {{0, 3, 2, 1}, {0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}, {4, 5, 6, 7}};
// Tet10 edges (cell_tet10.C:45-53): mid node 4+e joins these vertices.
const int kTetEdge[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};

inline double unit_uniform(std::mt19937_64 &g) { return (double)(g() >> 11) * 0x1.0p-53; }

// prepare_for_use -> renumber_nodes_and_elements: touched nodes get ids in
// first-touch order (elements in id order, local nodes in order); untouched
// nodes are dropped.
void renumber_first_touch(ReplicatedMesh &m) {
  const size_t nn = m._xyz.size() / 3;
  std::vector<int32_t> nid(nn, -1);
  int32_t next = 0;
  for (int32_t &c : m._conn) {
    if (nid[c] < 0) nid[c] = next++;
    c = nid[c];
  }
  std::vector<double> xyz(3 * (size_t)next);
  std::vector<uint8_t> bm(next);
  for (size_t n = 0; n < nn; ++n) {
    const int32_t t = nid[n];
    if (t < 0) continue;
    std::copy(&m._xyz[3 * n], &m._xyz[3 * n] + 3, &xyz[3 * (size_t)t]);
    bm[t] = m._bmask[n];
  }
  m._xyz.swap(xyz);
  m._bmask.swap(bm);
}

}  // namespace

namespace MeshTools {
namespace Generation {

void build_cube(ReplicatedMesh &mesh, unsigned nx, unsigned ny, unsigned nz, Real xmin, Real xmax, Real ymin,
                Real ymax, Real zmin, Real zmax, ElemType type) {
  if (nx == 0 || ny == 0 || nz == 0) libmesh_error_msg("build_cube needs nx, ny, nz > 0");
  if (type != HEX8 && type != HEX27 && type != TET4 && type != TET10)
    libmesh_error_msg("build_cube: element type " << (int)type << " is not supported here");
  const bool linear = (type == HEX8);
  const int f = linear ? 1 : 2;
  const int N[3] = {f * (int)nx, f * (int)ny, f * (int)nz};
  const Real lo[3] = {xmin, ymin, zmin}, span[3] = {xmax - xmin, ymax - ymin, zmax - zmin};
  mesh.clear();
  mesh._cells = {N[0], N[1], N[2]};
  const size_t nn = (size_t)(N[0] + 1) * (N[1] + 1) * (N[2] + 1);
  mesh._xyz.resize(3 * nn);
  mesh._bmask.resize(nn);
  // node lattice, x fastest; nodeset ids z0=0, y0=1, x1=2, y1=3, x0=4, z1=5
  size_t id = 0;
  for (int k = 0; k <= N[2]; ++k)
    for (int j = 0; j <= N[1]; ++j)
      for (int i = 0; i <= N[0]; ++i, ++id) {
        const int ijk[3] = {i, j, k};
        for (int c = 0; c < 3; ++c) mesh._xyz[3 * id + c] = (double)ijk[c] / (double)N[c] * span[c] + lo[c];
        uint8_t b = 0;
        if (k == 0) b |= 1u;
        if (j == 0) b |= 2u;
        if (i == N[0]) b |= 4u;
        if (j == N[1]) b |= 8u;
        if (i == 0) b |= 16u;
        if (k == N[2]) b |= 32u;
        mesh._bmask[id] = b;
      }
  auto lattice = [&](int i, int j, int k) { return (int32_t)(i + (N[0] + 1) * (j + (size_t)(N[1] + 1) * k)); };

  // hexes in k, j, i order
  const int hnpe = linear ? 8 : 27;
  std::vector<int32_t> hconn;
  std::vector<int8_t> hsides;
  hconn.reserve((size_t)nx * ny * nz * hnpe);
  for (unsigned k = 0; k < nz; ++k)
    for (unsigned j = 0; j < ny; ++j)
      for (unsigned i = 0; i < nx; ++i) {
        for (int a = 0; a < hnpe; ++a) {
          int o[3];
          for (int c = 0; c < 3; ++c) o[c] = linear ? kHex8Idx[c][a] : kHalfOffset[kHex27Idx[c][a]];
          hconn.push_back(lattice(f * i + o[0], f * j + o[1], f * k + o[2]));
        }
        int8_t s[6] = {-1, -1, -1, -1, -1, -1};
        if (k == 0) s[0] = 0;
        if (j == 0) s[1] = 1;
        if (i == nx - 1) s[2] = 2;
        if (j == ny - 1) s[3] = 3;
        if (i == 0) s[4] = 4;
        if (k == nz - 1) s[5] = 5;
        hsides.insert(hsides.end(), s, s + 6);
      }

  if (type == HEX8 || type == HEX27) {
    mesh._type = type;
    mesh._npe = hnpe;
    mesh._nsides = 6;
    mesh._conn.swap(hconn);
    mesh._sbid.swap(hsides);
    renumber_first_touch(mesh);
    return;
  }

  // Each HEX27 becomes 24 tets: per side (cycle v0..v3, centre c) and
  // t = 0..3 the tet (v_t, c, v_{t+1 mod 4}, hex centre 26), whose side 0 lies
  // on the hex side and inherits its boundary id (mesh_generation.C:1797-1850).
  const size_t nhex = (size_t)nx * ny * nz;
  std::vector<int32_t> tconn;
  std::vector<int8_t> tsides;
  tconn.reserve(nhex * 24 * 4);
  for (size_t h = 0; h < nhex; ++h) {
    const int32_t *hc = &hconn[h * 27];
    for (int s = 0; s < 6; ++s)
      for (int t = 0; t < 4; ++t) {
        tconn.push_back(hc[kSideCycle[s][t]]);
        tconn.push_back(hc[20 + s]);
        tconn.push_back(hc[kSideCycle[s][(t + 1) & 3]]);
        tconn.push_back(hc[26]);
        const int8_t ts[4] = {hsides[h * 6 + s], -1, -1, -1};
        tsides.insert(tsides.end(), ts, ts + 4);
      }
  }
  mesh._nsides = 4;
  mesh._sbid.swap(tsides);
  if (type == TET4) {
    mesh._type = TET4;
    mesh._npe = 4;
    mesh._conn.swap(tconn);
    renumber_first_touch(mesh);
    return;
  }
  // all_second_order: one node per distinct edge at the mean of its vertices
  std::map<std::pair<int32_t, int32_t>, int32_t> edge_node;
  std::vector<int32_t> qconn;
  const size_t ntet = tconn.size() / 4;
  qconn.reserve(ntet * 10);
  for (size_t e = 0; e < ntet; ++e) {
    const int32_t *v = &tconn[4 * e];
    qconn.insert(qconn.end(), v, v + 4);
    for (int ed = 0; ed < 6; ++ed) {
      const int32_t a = v[kTetEdge[ed][0]], b = v[kTetEdge[ed][1]];
      const auto key = std::make_pair(std::min(a, b), std::max(a, b));
      auto it = edge_node.find(key);
      if (it == edge_node.end()) {
        const int32_t nid = (int32_t)(mesh._xyz.size() / 3);
        for (int c = 0; c < 3; ++c) mesh._xyz.push_back((mesh._xyz[3 * (size_t)a + c] + mesh._xyz[3 * (size_t)b + c]) / 2.0);
        mesh._bmask.push_back(mesh._bmask[a] & mesh._bmask[b]);
        it = edge_node.emplace(key, nid).first;
      }
      qconn.push_back(it->second);
    }
  }
  mesh._type = TET10;
  mesh._npe = 10;
  mesh._conn.swap(qconn);
  renumber_first_touch(mesh);
}

}  // namespace Generation

namespace Modification {
void perturb_interior(ReplicatedMesh &mesh, uint64_t seed, Real frac) {
  std::mt19937_64 g(seed);
  const dof_id_type nn = mesh.n_nodes();
  for (dof_id_type n = 0; n < nn; ++n) {
    if (mesh._bmask[n]) continue;
    for (int c = 0; c < 3; ++c) {
      const double h = 1.0 / (double)mesh._cells[c];
      mesh._xyz[3 * (size_t)n + c] += frac * h * (2.0 * unit_uniform(g) - 1.0);
    }
  }
}
}  // namespace Modification

std::vector<Real> lognormal_coefficients(dof_id_type n_elem, uint64_t seed) {
  const double two_pi = 2.0 * 3.14159265358979323846264338327950288;
  std::mt19937_64 g(seed);
  std::vector<Real> k(n_elem);
  for (dof_id_type e = 0; e < n_elem; ++e) {
    const double u1 = 1.0 - unit_uniform(g), u2 = unit_uniform(g);
    k[e] = std::exp(std::sqrt(-2.0 * std::log(u1)) * std::cos(two_pi * u2));
  }
  return k;
}

void batch_recipe(uint64_t seed, int count, int nmin, int nmax, std::vector<int> &edge, std::vector<ElemType> &type) {
  static const ElemType kinds[4] = {HEX8, HEX27, TET4, TET10};
  std::mt19937_64 g(seed);
  edge.resize(count);
  type.resize(count);
  for (int p = 0; p < count; ++p) {
    const double u = unit_uniform(g);
    int n = (int)std::floor(nmin * std::pow((double)nmax / nmin, u) + 0.5);
    edge[p] = std::min(nmax, std::max(nmin, n));
    type[p] = kinds[std::min(3, (int)(4.0 * unit_uniform(g)))];
  }
}

}  // namespace MeshTools
}  // namespace libMesh
