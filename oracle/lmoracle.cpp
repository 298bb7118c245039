// lmoracle.cpp -- CPU oracle restating libMesh's FP64 assembly + Jacobi-CG path.
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py as the checker; the product never links it.
//
// Every block cites the libMesh source it restates (paths relative to the
// libMesh tree).  Pinning status of each function is listed in the header of
// tests/test_oracle_pins.py.
#include "lmoracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <random>
#include <set>
#include <utility>
#include <vector>

namespace {

using i64 = int64_t;

// ---------------------------------------------------------------------------
// Mesh
// ---------------------------------------------------------------------------
struct Mesh {
  int type = 0, npe = 0, nsides = 0;
  i64 n_nodes = 0, n_elem = 0;
  int cells[3] = {0, 0, 0};      // lattice cells per axis (node spacing 1/cells)
  std::vector<double> xyz;       // 3 * n_nodes
  std::vector<int32_t> conn;     // npe * n_elem
  std::vector<uint8_t> bmask;    // nodeset boundary ids (bit b = id b)
  std::vector<int8_t> side_bid;  // nsides * n_elem, -1 = no boundary id
};

// Hex27::side_nodes_map (src/geom/cell_hex27.C:37-45); the first four entries of
// each row are Hex8::side_nodes_map.
const int kHexSide[6][9] = {{0, 3, 2, 1, 11, 10, 9, 8, 20},  {0, 1, 5, 4, 8, 13, 16, 12, 21},
                            {1, 2, 6, 5, 9, 14, 17, 13, 22}, {2, 3, 7, 6, 10, 15, 18, 14, 23},
                            {3, 0, 4, 7, 11, 12, 19, 15, 24}, {4, 5, 6, 7, 16, 17, 18, 19, 25}};
// Tet10::side_nodes_map (src/geom/cell_tet10.C:37-43); first three = Tet4.
const int kTetSide[4][6] = {{0, 2, 1, 6, 5, 4}, {0, 1, 3, 4, 8, 7}, {1, 2, 3, 5, 9, 8}, {2, 0, 3, 6, 7, 9}};
// Tet10::edge_nodes_map (src/geom/cell_tet10.C:45-53): mid node 4+e on edge e.
const int kTetEdge[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};

int nodes_per(int t) {
  switch (t) {
    case ORC_HEX8: return 8;
    case ORC_HEX27: return 27;
    case ORC_TET4: return 4;
    case ORC_TET10: return 10;
  }
  return 0;
}
int sides_per(int t) { return (t == ORC_HEX8 || t == ORC_HEX27) ? 6 : 4; }
int side_nodes(int t) { return t == ORC_HEX8 ? 4 : t == ORC_HEX27 ? 9 : t == ORC_TET4 ? 3 : 6; }
int side_node(int t, int s, int k) {
  return (t == ORC_HEX8 || t == ORC_HEX27) ? kHexSide[s][k] : kTetSide[s][k];
}

// ReplicatedMesh::renumber_nodes_and_elements (src/mesh/replicated_mesh.C:700-830):
// nodes are renumbered by first touch in element order and local node order;
// nodes no element touches are deleted.  (The swap-based implementation there
// assigns exactly the first-touch ids to touched nodes.)
void renumber(Mesh &m) {
  std::vector<int32_t> nid(m.n_nodes, -1);
  int32_t next = 0;
  for (i64 e = 0; e < m.n_elem; ++e)
    for (int a = 0; a < m.npe; ++a) {
      int32_t &t = nid[m.conn[e * m.npe + a]];
      if (t < 0) t = next++;
    }
  std::vector<double> xyz(3 * (size_t)next);
  std::vector<uint8_t> bm(next);
  for (i64 n = 0; n < m.n_nodes; ++n)
    if (nid[n] >= 0) {
      for (int c = 0; c < 3; ++c) xyz[3 * nid[n] + c] = m.xyz[3 * n + c];
      bm[nid[n]] = m.bmask[n];
    }
  for (auto &c : m.conn) c = nid[c];
  m.xyz.swap(xyz);
  m.bmask.swap(bm);
  m.n_nodes = next;
}

// MeshTools::Generation::build_cube, 3D branch (src/mesh/mesh_generation.C:1160-1940)
// on [0,1]^3: node lattice k-outer, j, i-inner (:1329-1350); nodeset ids
// z0=0, z1=5, y0=1, y1=3, x0=4, x1=2; HEX8 elements (:1398-1432); HEX27
// elements (:1560-1625); TET4 = 24 tets per HEX27 (:1797-1850); TET10 via
// all_second_order (:1856).
Mesh *build_cube(int nx, int ny, int nz, int type) {
  Mesh *m = new Mesh;
  m->type = type;
  const bool lin = (type == ORC_HEX8);
  const int f = lin ? 1 : 2;
  const int Nx = f * nx, Ny = f * ny, Nz = f * nz;
  m->cells[0] = Nx;
  m->cells[1] = Ny;
  m->cells[2] = Nz;
  const i64 nn = (i64)(Nx + 1) * (Ny + 1) * (Nz + 1);
  m->xyz.resize(3 * nn);
  m->bmask.assign(nn, 0);
  i64 id = 0;
  for (int k = 0; k <= Nz; ++k)
    for (int j = 0; j <= Ny; ++j)
      for (int i = 0; i <= Nx; ++i, ++id) {
        // Point(i/Nx, j/Ny, k/Nz), then scaled by (max-min)=1 and shifted by min=0.
        m->xyz[3 * id + 0] = (double)i / (double)Nx * 1.0 + 0.0;
        m->xyz[3 * id + 1] = (double)j / (double)Ny * 1.0 + 0.0;
        m->xyz[3 * id + 2] = (double)k / (double)Nz * 1.0 + 0.0;
        uint8_t b = 0;
        if (k == 0) b |= 1u << 0;
        if (k == Nz) b |= 1u << 5;
        if (j == 0) b |= 1u << 1;
        if (j == Ny) b |= 1u << 3;
        if (i == 0) b |= 1u << 4;
        if (i == Nx) b |= 1u << 2;
        m->bmask[id] = b;
      }
  m->n_nodes = nn;
  auto idx = [&](int i, int j, int k) -> int32_t { return (int32_t)(i + (Nx + 1) * (j + (i64)(Ny + 1) * k)); };

  std::vector<int32_t> conn;
  std::vector<int8_t> sbid;
  const int hnpe = lin ? 8 : 27;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        int32_t c[27];
        if (lin) {
          c[0] = idx(i, j, k);         c[1] = idx(i + 1, j, k);
          c[2] = idx(i + 1, j + 1, k); c[3] = idx(i, j + 1, k);
          c[4] = idx(i, j, k + 1);     c[5] = idx(i + 1, j, k + 1);
          c[6] = idx(i + 1, j + 1, k + 1); c[7] = idx(i, j + 1, k + 1);
        } else {
          const int I = 2 * i, J = 2 * j, K = 2 * k;
          c[0] = idx(I, J, K);          c[1] = idx(I + 2, J, K);
          c[2] = idx(I + 2, J + 2, K);  c[3] = idx(I, J + 2, K);
          c[4] = idx(I, J, K + 2);      c[5] = idx(I + 2, J, K + 2);
          c[6] = idx(I + 2, J + 2, K + 2); c[7] = idx(I, J + 2, K + 2);
          c[8] = idx(I + 1, J, K);      c[9] = idx(I + 2, J + 1, K);
          c[10] = idx(I + 1, J + 2, K); c[11] = idx(I, J + 1, K);
          c[12] = idx(I, J, K + 1);     c[13] = idx(I + 2, J, K + 1);
          c[14] = idx(I + 2, J + 2, K + 1); c[15] = idx(I, J + 2, K + 1);
          c[16] = idx(I + 1, J, K + 2); c[17] = idx(I + 2, J + 1, K + 2);
          c[18] = idx(I + 1, J + 2, K + 2); c[19] = idx(I, J + 1, K + 2);
          c[20] = idx(I + 1, J + 1, K); c[21] = idx(I + 1, J, K + 1);
          c[22] = idx(I + 2, J + 1, K + 1); c[23] = idx(I + 1, J + 2, K + 1);
          c[24] = idx(I, J + 1, K + 1); c[25] = idx(I + 1, J + 1, K + 2);
          c[26] = idx(I + 1, J + 1, K + 1);
        }
        conn.insert(conn.end(), c, c + hnpe);
        int8_t sb[6] = {-1, -1, -1, -1, -1, -1};
        if (k == 0) sb[0] = 0;
        if (k == nz - 1) sb[5] = 5;
        if (j == 0) sb[1] = 1;
        if (j == ny - 1) sb[3] = 3;
        if (i == 0) sb[4] = 4;
        if (i == nx - 1) sb[2] = 2;
        sbid.insert(sbid.end(), sb, sb + 6);
      }

  if (type == ORC_HEX8 || type == ORC_HEX27) {
    m->npe = hnpe;
    m->nsides = 6;
    m->conn.swap(conn);
    m->side_bid.swap(sbid);
    m->n_elem = (i64)nx * ny * nz;
    renumber(*m);
    return m;
  }

  // 24 sub-tets per HEX27: for each side s (QUAD9 in side_nodes_map order) and
  // sub_tet t: (side[t], side[8], side[(t+1)%4], hex node 26); the tet's side 0
  // inherits the hex side's boundary id (mesh_generation.C:1808-1840).
  const i64 nhex = (i64)nx * ny * nz;
  std::vector<int32_t> tconn;
  std::vector<int8_t> tsb;
  tconn.reserve(nhex * 24 * 4);
  for (i64 h = 0; h < nhex; ++h) {
    const int32_t *c = &conn[h * 27];
    for (int s = 0; s < 6; ++s) {
      const int8_t b = sbid[h * 6 + s];
      for (int t = 0; t < 4; ++t) {
        tconn.push_back(c[kHexSide[s][t]]);
        tconn.push_back(c[kHexSide[s][8]]);
        tconn.push_back(c[kHexSide[s][t == 3 ? 0 : t + 1]]);
        tconn.push_back(c[26]);
        int8_t sb[4] = {b, -1, -1, -1};
        tsb.insert(tsb.end(), sb, sb + 4);
      }
    }
  }
  m->n_elem = nhex * 24;
  m->nsides = 4;
  if (type == ORC_TET4) {
    m->npe = 4;
    m->conn.swap(tconn);
    m->side_bid.swap(tsb);
    renumber(*m);
    return m;
  }
  // all_second_order (src/mesh/unstructured_mesh.C:1732ff): one new node per
  // distinct edge (keyed by the sorted vertex pair) at the average of its two
  // vertices; the mesh is then re-prepared, i.e. renumbered by first touch.
  std::map<std::pair<int32_t, int32_t>, int32_t> mid;
  std::vector<int32_t> qconn;
  qconn.reserve(m->n_elem * 10);
  for (i64 e = 0; e < m->n_elem; ++e) {
    const int32_t *v = &tconn[e * 4];
    for (int a = 0; a < 4; ++a) qconn.push_back(v[a]);
    for (int ed = 0; ed < 6; ++ed) {
      int32_t a = v[kTetEdge[ed][0]], b = v[kTetEdge[ed][1]];
      auto key = std::make_pair(std::min(a, b), std::max(a, b));
      auto it = mid.find(key);
      if (it == mid.end()) {
        const int32_t nid = (int32_t)m->n_nodes++;
        for (int cc = 0; cc < 3; ++cc) {
          double p = m->xyz[3 * (size_t)a + cc];
          p += m->xyz[3 * (size_t)b + cc];
          m->xyz.push_back(p / 2.0);
        }
        m->bmask.push_back(m->bmask[a] & m->bmask[b]);
        it = mid.emplace(key, nid).first;
      }
      qconn.push_back(it->second);
    }
  }
  m->npe = 10;
  m->conn.swap(qconn);
  m->side_bid.swap(tsb);
  renumber(*m);
  return m;
}

// ---------------------------------------------------------------------------
// Quadrature (src/quadrature/quadrature_gauss_1D.C:30-80, quadrature_gauss_3D.C,
// quadrature.C:283-305)
// ---------------------------------------------------------------------------
void qgauss_1d(int np, double *x, double *w) {
  if (np == 1) {
    x[0] = 0.;
    w[0] = 2.;
  } else if (np == 2) {
    x[0] = -std::sqrt(3.0) / 3;
    x[1] = -x[0];
    w[0] = 1.;
    w[1] = w[0];
  } else {  // np == 3
    x[0] = -7.7459666924148337703585307995648e-01;
    x[1] = 0.;
    x[2] = -x[0];
    w[0] = 5.0 / 9;
    w[1] = 8.0 / 9;
    w[2] = w[0];
  }
}

struct QRule {
  std::vector<double> p;  // 3 * n
  std::vector<double> w;
  int n = 0;
};

// QBase::tensor_product_hex: q = i + np*(j + np*k), w = w_i * w_j * w_k.
QRule qgauss_hex(int np) {
  double x[3], w[3];
  qgauss_1d(np, x, w);
  QRule r;
  for (int k = 0; k < np; ++k)
    for (int j = 0; j < np; ++j)
      for (int i = 0; i < np; ++i) {
        r.p.push_back(x[i]);
        r.p.push_back(x[j]);
        r.p.push_back(x[k]);
        r.w.push_back(w[i] * w[j] * w[k]);
      }
  r.n = np * np * np;
  return r;
}

// QGauss on TET (quadrature_gauss_3D.C:55-260): THIRD = 5-point rule with a
// negative weight (allow_rules_with_negative_weights defaults to true,
// quadrature.C:29); FIFTH = 14-point Walkington rule.
QRule qgauss_tet(int order) {
  QRule r;
  auto add = [&](double a, double b, double c, double w) {
    r.p.push_back(a);
    r.p.push_back(b);
    r.p.push_back(c);
    r.w.push_back(w);
  };
  if (order <= 1) {
    add(.25, .25, .25, 1.0 / 6);
  } else if (order == 2) {
    const double b = 0.25 * (1 - std::sqrt(5.0) / 5);
    const double a = 1 - 3 * b;
    add(a, b, b, 1.0 / 24);
    add(b, a, b, 1.0 / 24);
    add(b, b, a, 1.0 / 24);
    add(b, b, b, 1.0 / 24);
  } else if (order == 3) {
    add(.25, .25, .25, -2.0 / 15);
    add(.5, 1.0 / 6, 1.0 / 6, .075);
    add(1.0 / 6, .5, 1.0 / 6, .075);
    add(1.0 / 6, 1.0 / 6, .5, .075);
    add(1.0 / 6, 1.0 / 6, 1.0 / 6, .075);
  } else {
    const double a[3] = {0.31088591926330060980, 0.092735250310891226402, 0.045503704125649649492};
    const double wt[3] = {0.018781320953002641800, 0.012248840519393658257, 0.0070910034628469110730};
    for (int i = 0; i < 2; ++i) {
      const double b = 1. - 3. * a[i];
      add(a[i], a[i], a[i], wt[i]);
      add(a[i], b, a[i], wt[i]);
      add(b, a[i], a[i], wt[i]);
      add(a[i], a[i], b, wt[i]);
    }
    const double b = 0.5 * (1. - 2. * a[2]);
    add(b, b, a[2], wt[2]);
    add(b, a[2], a[2], wt[2]);
    add(a[2], a[2], b, wt[2]);
    add(a[2], b, a[2], wt[2]);
    add(b, a[2], b, wt[2]);
    add(a[2], b, b, wt[2]);
  }
  r.n = (int)r.w.size();
  return r;
}

// ---------------------------------------------------------------------------
// Lagrange shapes (src/fe/fe_lagrange_shape_3D.C:73-415, 700-1010, 1486-2300;
// include/fe/fe_lagrange_shape_1D.h:32-140)
// ---------------------------------------------------------------------------
inline double l1(int i, double x) { return i == 0 ? .5 * (1. - x) : .5 * (1. + x); }
inline double dl1(int i) { return i == 0 ? -.5 : .5; }
inline double l2(int i, double x) { return i == 0 ? .5 * x * (x - 1.) : i == 1 ? .5 * x * (x + 1) : (1. - x * x); }
inline double dl2(int i, double x) { return i == 0 ? x - .5 : i == 1 ? x + .5 : -2. * x; }
const int kH8[3][8] = {{0, 1, 1, 0, 0, 1, 1, 0}, {0, 0, 1, 1, 0, 0, 1, 1}, {0, 0, 0, 0, 1, 1, 1, 1}};
const int kH27[3][27] = {
    {0, 1, 1, 0, 0, 1, 1, 0, 2, 1, 2, 0, 0, 1, 1, 0, 2, 1, 2, 0, 2, 2, 1, 2, 0, 2, 2},
    {0, 0, 1, 1, 0, 0, 1, 1, 0, 2, 1, 2, 0, 0, 1, 1, 0, 2, 1, 2, 2, 0, 2, 1, 2, 2, 2},
    {0, 0, 0, 0, 1, 1, 1, 1, 0, 0, 0, 0, 2, 2, 2, 2, 1, 1, 1, 1, 0, 2, 2, 2, 2, 1, 2}};

// phi[i*nq+q], dphi[(i*nq+q)*3+c]
int shapes(int type, int nq, const double *pts, std::vector<double> &phi, std::vector<double> &dphi) {
  const int n = nodes_per(type);
  phi.assign((size_t)n * nq, 0.);
  dphi.assign((size_t)n * nq * 3, 0.);
  for (int q = 0; q < nq; ++q) {
    const double xi = pts[3 * q], eta = pts[3 * q + 1], zeta = pts[3 * q + 2];
    for (int i = 0; i < n; ++i) {
      double *d = &dphi[((size_t)i * nq + q) * 3];
      double &p = phi[(size_t)i * nq + q];
      if (type == ORC_HEX8) {
        const int a = kH8[0][i], b = kH8[1][i], c = kH8[2][i];
        p = l1(a, xi) * l1(b, eta) * l1(c, zeta);
        d[0] = dl1(a) * l1(b, eta) * l1(c, zeta);
        d[1] = l1(a, xi) * dl1(b) * l1(c, zeta);
        d[2] = l1(a, xi) * l1(b, eta) * dl1(c);
      } else if (type == ORC_HEX27) {
        const int a = kH27[0][i], b = kH27[1][i], c = kH27[2][i];
        p = l2(a, xi) * l2(b, eta) * l2(c, zeta);
        d[0] = dl2(a, xi) * l2(b, eta) * l2(c, zeta);
        d[1] = l2(a, xi) * dl2(b, eta) * l2(c, zeta);
        d[2] = l2(a, xi) * l2(b, eta) * dl2(c, zeta);
      } else {
        // Area coordinates zeta0 = 1 - xi - eta - zeta (fe_lagrange_shape_3D.C:715-740,
        // 957-1005); derivatives are the exact partials of those expressions.
        const double z1 = xi, z2 = eta, z3 = zeta, z0 = 1. - z1 - z2 - z3;
        const double g[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        const double z[4] = {z0, z1, z2, z3};
        if (type == ORC_TET4) {
          p = z[i];
          for (int cc = 0; cc < 3; ++cc) d[cc] = g[i][cc];
        } else if (i < 4) {
          p = z[i] * (2. * z[i] - 1.);
          for (int cc = 0; cc < 3; ++cc) d[cc] = (4. * z[i] - 1.) * g[i][cc];
        } else {
          const int a = kTetEdge[i - 4][0], b = kTetEdge[i - 4][1];
          // node 6 is 4*zeta2*zeta0 in the reference; the product is symmetric.
          p = 4. * z[a] * z[b];
          for (int cc = 0; cc < 3; ++cc) d[cc] = 4. * (g[a][cc] * z[b] + z[a] * g[b][cc]);
        }
      }
    }
  }
  return n;
}

// ---------------------------------------------------------------------------
// Affine-map test: Hex8::has_affine_map (src/geom/cell_hex8.C:121-135),
// Hex27::has_affine_map (src/geom/cell_hex27.C:125-160); Tet4 is always affine;
// Tet10 is affine when every mid-edge node is its edge's midpoint.
// relative_fuzzy_equals: |a-b|_1 <= tol*(|a|_1+|b|_1), tol = affine_tol = 1e-12
// (include/utils/fuzzy_equals.h:109, include/geom/elem.h:2070).
// ---------------------------------------------------------------------------
struct P3 {
  double x, y, z;
};
inline P3 sub(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline P3 half(P3 a) { return {a.x / 2, a.y / 2, a.z / 2}; }
inline double l1n(P3 a) { return std::abs(a.x) + std::abs(a.y) + std::abs(a.z); }
inline bool rfe(P3 a, P3 b) {
  const double tol = 1e-12;
  return l1n(sub(a, b)) <= tol * (l1n(a) + l1n(b));
}

bool has_affine_map(int type, const P3 *p) {
  if (type == ORC_TET4) return true;
  if (type == ORC_TET10) {
    for (int e = 0; e < 6; ++e) {
      P3 a = p[kTetEdge[e][0]], b = p[kTetEdge[e][1]];
      P3 m = {(a.x + b.x) / 2, (a.y + b.y) / 2, (a.z + b.z) / 2};
      if (!rfe(m, p[4 + e])) return false;
    }
    return true;
  }
  P3 v = sub(p[1], p[0]);
  if (!rfe(v, sub(p[2], p[3])) || !rfe(v, sub(p[5], p[4])) || !rfe(v, sub(p[6], p[7]))) return false;
  if (type == ORC_HEX8) {
    v = sub(p[4], p[0]);
    return rfe(v, sub(p[7], p[3]));
  }
  v = half(v);
  if (!rfe(v, sub(p[8], p[0])) || !rfe(v, sub(p[10], p[3])) || !rfe(v, sub(p[16], p[4])) ||
      !rfe(v, sub(p[18], p[7])) || !rfe(v, sub(p[20], p[11])) || !rfe(v, sub(p[21], p[12])) ||
      !rfe(v, sub(p[23], p[15])) || !rfe(v, sub(p[25], p[19])) || !rfe(v, sub(p[26], p[24])))
    return false;
  v = sub(p[4], p[0]);
  if (!rfe(v, sub(p[7], p[3]))) return false;
  v = half(v);
  if (!rfe(v, sub(p[12], p[0])) || !rfe(v, sub(p[13], p[1])) || !rfe(v, sub(p[14], p[2])) ||
      !rfe(v, sub(p[15], p[3])) || !rfe(v, sub(p[22], p[9])) || !rfe(v, sub(p[24], p[11])))
    return false;
  v = half(sub(p[3], p[0]));
  if (!rfe(v, sub(p[11], p[0])) || !rfe(v, sub(p[9], p[1])) || !rfe(v, sub(p[17], p[5])) ||
      !rfe(v, sub(p[19], p[4])))
    return false;
  return true;
}

// ---------------------------------------------------------------------------
// FE evaluator: FE<3,LAGRANGE>::reinit (src/fe/fe.C:197-368) with
// FEMap::compute_map (fe_map.C:1323-1372), compute_single_point_map 3D
// (fe_map.C:958-1090), compute_affine_map (fe_map.C:1167-1246) and
// H1FETransformation::map_dphi 3D (src/fe/h1_fe_transformation.C:64ff).
// ---------------------------------------------------------------------------
struct FEValues {
  int type = 0, n = 0, nq = 0;
  QRule q;
  std::vector<double> phi, dref;       // reference tables (cached per type, fe.C:246-283)
  std::vector<double> JxW, xyz, dphi;  // per element: nq, 3*nq, n*nq*3
  std::vector<double> jac;

  void init(int t, int npts_or_order) {
    type = t;
    q = (t == ORC_HEX8 || t == ORC_HEX27) ? qgauss_hex(npts_or_order) : qgauss_tet(npts_or_order);
    nq = q.n;
    n = shapes(t, nq, q.p.data(), phi, dref);
    JxW.resize(nq);
    jac.resize(nq);
    xyz.resize(3 * nq);
    dphi.resize((size_t)n * nq * 3);
  }

  // Returns false on a degenerate map (det J <= jacobian_tolerance = 0, fe_map.C:467-505).
  bool reinit(const Mesh &m, i64 e) {
    P3 pt[27];
    for (int a = 0; a < n; ++a) {
      const int32_t nd = m.conn[e * m.npe + a];
      pt[a] = {m.xyz[3 * (size_t)nd], m.xyz[3 * (size_t)nd + 1], m.xyz[3 * (size_t)nd + 2]};
    }
    const bool affine = has_affine_map(type, pt);
    double inv[9];
    for (int p = 0; p < nq; ++p) {
      // xyz at every qp (compute_affine_map computes xyz for all p as well)
      double X = 0, Y = 0, Z = 0;
      for (int a = 0; a < n; ++a) {
        const double s = phi[(size_t)a * nq + p];
        X += pt[a].x * s;
        Y += pt[a].y * s;
        Z += pt[a].z * s;
      }
      xyz[3 * p] = X;
      xyz[3 * p + 1] = Y;
      xyz[3 * p + 2] = Z;
      if (p == 0 || !affine) {
        double dxi[3] = {0, 0, 0}, deta[3] = {0, 0, 0}, dzeta[3] = {0, 0, 0};
        for (int a = 0; a < n; ++a) {  // add_scaled(elem_point, dphidxi_map[i][p]) ...
          const double *d = &dref[((size_t)a * nq + p) * 3];
          dxi[0] += pt[a].x * d[0];
          dxi[1] += pt[a].y * d[0];
          dxi[2] += pt[a].z * d[0];
          deta[0] += pt[a].x * d[1];
          deta[1] += pt[a].y * d[1];
          deta[2] += pt[a].z * d[1];
          dzeta[0] += pt[a].x * d[2];
          dzeta[1] += pt[a].y * d[2];
          dzeta[2] += pt[a].z * d[2];
        }
        const double dx_dxi = dxi[0], dy_dxi = dxi[1], dz_dxi = dxi[2];
        const double dx_deta = deta[0], dy_deta = deta[1], dz_deta = deta[2];
        const double dx_dzeta = dzeta[0], dy_dzeta = dzeta[1], dz_dzeta = dzeta[2];
        const double J = (dx_dxi * (dy_deta * dz_dzeta - dz_deta * dy_dzeta) +
                          dy_dxi * (dz_deta * dx_dzeta - dx_deta * dz_dzeta) +
                          dz_dxi * (dx_deta * dy_dzeta - dy_deta * dx_dzeta));
        if (!(J > 0.0)) return false;
        jac[p] = J;
        JxW[p] = J * q.w[p];
        const double ij = 1. / J;
        inv[0] = (dy_deta * dz_dzeta - dz_deta * dy_dzeta) * ij;  // dxidx
        inv[1] = (dz_deta * dx_dzeta - dx_deta * dz_dzeta) * ij;  // dxidy
        inv[2] = (dx_deta * dy_dzeta - dy_deta * dx_dzeta) * ij;  // dxidz
        inv[3] = (dz_dxi * dy_dzeta - dy_dxi * dz_dzeta) * ij;    // detadx
        inv[4] = (dx_dxi * dz_dzeta - dz_dxi * dx_dzeta) * ij;    // detady
        inv[5] = (dy_dxi * dx_dzeta - dx_dxi * dy_dzeta) * ij;    // detadz
        inv[6] = (dy_dxi * dz_deta - dz_dxi * dy_deta) * ij;      // dzetadx
        inv[7] = (dz_dxi * dx_deta - dx_dxi * dz_deta) * ij;      // dzetady
        inv[8] = (dx_dxi * dy_deta - dy_dxi * dx_deta) * ij;      // dzetadz
      } else {
        jac[p] = jac[0];
        JxW[p] = JxW[0] / q.w[0] * q.w[p];
      }
      for (int a = 0; a < n; ++a) {
        const double *d = &dref[((size_t)a * nq + p) * 3];
        double *o = &dphi[((size_t)a * nq + p) * 3];
        o[0] = (d[0] * inv[0] + d[1] * inv[3] + d[2] * inv[6]);
        o[1] = (d[0] * inv[1] + d[1] * inv[4] + d[2] * inv[7]);
        o[2] = (d[0] * inv[2] + d[1] * inv[5] + d[2] * inv[8]);
      }
    }
    return true;
  }
  double ph(int a, int p) const { return phi[(size_t)a * nq + p]; }
  const double *dp(int a, int p) const { return &dphi[((size_t)a * nq + p) * 3]; }
};

inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

const double kPi = 3.14159265358979323846264338327950288;  // libMesh::pi

double forcing(int f_kind, const double *x) {
  if (f_kind == 0) return 1.0;
  return 3. * kPi * kPi * std::sin(kPi * x[0]) * std::sin(kPi * x[1]) * std::sin(kPi * x[2]);
}

// ---------------------------------------------------------------------------
// System: DoF map, constraint set, sparsity
// ---------------------------------------------------------------------------
using ConstraintRow = std::map<int32_t, double>;  // DofConstraintRow

struct System {
  const Mesh *m = nullptr;
  int nv = 1, dpe = 0;
  i64 ndofs = 0;
  std::vector<int32_t> edofs;   // n_elem * dpe
  std::vector<uint8_t> cons;    // per dof
  std::vector<int32_t> rp, ci;  // CSR pattern
  // DofMap::_dof_constraints and _primal_constraint_values; Dirichlet DoFs hold
  // empty rows.  `general` is set once a row with entries or a nonzero value
  // exists: assembly then follows the dense constraint-matrix path.
  std::map<int32_t, ConstraintRow> crows;
  std::map<int32_t, double> cvals;
  bool general = false;
  bool asymmetric = true;  // asymmetric_constraint_rows (dof_map.h default true)
  bool is_constrained(int32_t d) const { return crows.count(d) != 0; }
};

// distribute_local_dofs_var_major (src/base/dof_map.C:1417-1530): nodes get a
// variable-group base in first-touch order over active elements (element id
// order, local node order); DofObject::dof_number = base + var*n_comp + comp
// (include/base/dof_object.h:1055) with n_comp = 1 for Lagrange.  Because the
// mesh was renumbered by first touch over the same traversal, base = nv*node.
// dof_indices (dof_map.C:2201-2316) lists variable-major: all nodes of var 0,
// then var 1, ...
void build_dofs(System &s) {
  const Mesh &m = *s.m;
  s.dpe = s.nv * m.npe;
  std::vector<int32_t> base(m.n_nodes, -1);
  int32_t next = 0;
  for (i64 e = 0; e < m.n_elem; ++e)
    for (int a = 0; a < m.npe; ++a) {
      int32_t &b = base[m.conn[e * m.npe + a]];
      if (b < 0) {
        b = next;
        next += s.nv;
      }
    }
  s.ndofs = next;
  s.edofs.resize((size_t)m.n_elem * s.dpe);
  for (i64 e = 0; e < m.n_elem; ++e)
    for (int v = 0; v < s.nv; ++v)
      for (int a = 0; a < m.npe; ++a)
        s.edofs[(size_t)e * s.dpe + v * m.npe + a] = base[m.conn[e * m.npe + a]] + v;
}

// DirichletBoundary over boundary ids: every DoF of every node on an element
// side carrying one of the ids is constrained with an empty constraint row and
// value 0 (ConstrainDirichlet::apply_lagrange_dirichlet_impl,
// src/base/dof_map_constraints.C:592-740).
void build_constraints(System &s, unsigned mask) {
  const Mesh &m = *s.m;
  s.cons.assign(s.ndofs, 0);
  const int sn = side_nodes(m.type);
  for (i64 e = 0; e < m.n_elem; ++e)
    for (int sd = 0; sd < m.nsides; ++sd) {
      const int b = m.side_bid[e * m.nsides + sd];
      if (b < 0 || !(mask & (1u << b))) continue;
      for (int k = 0; k < sn; ++k) {
        const int a = side_node(m.type, sd, k);
        for (int v = 0; v < s.nv; ++v) s.cons[s.edofs[(size_t)e * s.dpe + v * m.npe + a]] = 1;
      }
    }
  s.crows.clear();
  s.cvals.clear();
  for (i64 d = 0; d < s.ndofs; ++d)
    if (s.cons[d]) s.crows[(int32_t)d];  // empty constraint row, value 0
}

// SparsityPattern::Build (src/base/sparsity_pattern.C:96-365): each row is the
// sorted, unique union of the DoFs of every element containing the row's DoF
// (DefaultCoupling with no neighbour layers; all variables coupled).
void build_sparsity(System &s) {
  const Mesh &m = *s.m;
  std::vector<i64> cnt(m.n_nodes + 1, 0);
  for (i64 e = 0; e < m.n_elem; ++e)
    for (int a = 0; a < m.npe; ++a) cnt[m.conn[e * m.npe + a] + 1]++;
  for (i64 n = 0; n < m.n_nodes; ++n) cnt[n + 1] += cnt[n];
  std::vector<int32_t> adj(cnt[m.n_nodes]);
  std::vector<i64> pos(cnt.begin(), cnt.end() - 1);
  for (i64 e = 0; e < m.n_elem; ++e)
    for (int a = 0; a < m.npe; ++a) adj[pos[m.conn[e * m.npe + a]]++] = (int32_t)e;
  // node-level pattern
  std::vector<std::vector<int32_t>> nrow(m.n_nodes);
#pragma omp parallel for schedule(dynamic, 1024)
  for (i64 n = 0; n < m.n_nodes; ++n) {
    std::vector<int32_t> &r = nrow[n];
    for (i64 k = cnt[n]; k < cnt[n + 1]; ++k) {
      const i64 e = adj[k];
      r.insert(r.end(), &m.conn[e * m.npe], &m.conn[e * m.npe] + m.npe);
    }
    std::sort(r.begin(), r.end());
    r.erase(std::unique(r.begin(), r.end()), r.end());
  }
  // node base -> dof rows; node ids equal dof bases / nv
  s.rp.assign(s.ndofs + 1, 0);
  for (i64 n = 0; n < m.n_nodes; ++n)
    for (int v = 0; v < s.nv; ++v) s.rp[(i64)s.nv * n + v + 1] = (int32_t)(nrow[n].size() * s.nv);
  for (i64 r = 0; r < s.ndofs; ++r) s.rp[r + 1] += s.rp[r];
  s.ci.resize(s.rp[s.ndofs]);
#pragma omp parallel for schedule(dynamic, 1024)
  for (i64 n = 0; n < m.n_nodes; ++n)
    for (int v = 0; v < s.nv; ++v) {
      int32_t *o = &s.ci[s.rp[(i64)s.nv * n + v]];
      for (int32_t c : nrow[n])
        for (int w = 0; w < s.nv; ++w) *o++ = s.nv * c + w;
    }
}

// DofMap::heterogeneously_constrain_element_matrix_and_vector
// (src/base/dof_map_constraints.C:2518-2616) for empty constraint rows with value
// 0: C = diag(free), so C^T K C keeps K(i,j) for free i,j and zeroes the rest,
// C^T (F - K H) keeps F(i) for free i; then each constrained row is zeroed with
// K(i,i) = 1 and rhs(i) = 0.  The elementwise mask below is that product
// (every term of the dense triple product is either the entry itself or an
// exact zero).
void constrain(const System &s, i64 e, double *K, double *F) {
  const int n = s.dpe;
  const int32_t *d = &s.edofs[(size_t)e * n];
  bool any = false;
  for (int i = 0; i < n; ++i) any |= s.cons[d[i]] != 0;
  if (!any) return;  // early return when no DoF is constrained (:3472, :3584)
  for (int i = 0; i < n; ++i) {
    const bool ci = s.cons[d[i]];
    for (int j = 0; j < n; ++j)
      if (ci || s.cons[d[j]]) K[i * n + j] = 0.;
    if (ci) {
      K[i * n + i] = 1.;
      if (F) F[i] = 0.;
    }
  }
}

// ---------------------------------------------------------------------------
// General constraint rows
// ---------------------------------------------------------------------------
// DofMap::add_constraint_row (src/base/dof_map_constraints.C:2180-2208): store
// the row and its value (insert_or_assign), refusing a constrained DoF when
// forbid_constraint_overwrite is set.
int add_constraint_row(System &s, int32_t dof, const ConstraintRow &row, double rhs, bool forbid) {
  if (forbid && s.is_constrained(dof)) return -1;
  s.crows[dof] = row;
  s.cvals[dof] = rhs;
  return 0;
}

// DofMap::check_for_constraint_loops (dof_map_constraints.C:4509-4597): the
// expansion of process_constraints on a copy of the rows, failing when a row
// being substituted refers to the DoF it constrains.
bool constraint_loop(const System &s) {
  std::map<int32_t, ConstraintRow> rows = s.crows;
  std::set<int32_t> todo;
  for (const auto &r : rows) todo.insert(r.first);
  while (!todo.empty())
    for (auto i = todo.begin(); i != todo.end();) {
      ConstraintRow &row = rows[*i];
      std::vector<int32_t> expand;
      for (const auto &it : row)
        if (it.first != *i && s.is_constrained(it.first)) {
          todo.insert(it.first);
          expand.push_back(it.first);
        }
      for (int32_t x : expand) {
        const double c = row[x];
        for (const auto &it : rows.at(x)) {
          if (it.first == x) return true;  // "Constraint loop detected"
          row[it.first] += it.second * c;
        }
        row.erase(x);
      }
      if (expand.empty())
        i = todo.erase(i);
      else
        ++i;
    }
  return false;
}

// DofMap::process_constraints (dof_map_constraints.C:4341-4497): substitute
// constrained DoFs inside every row (and their values into the row's value)
// until rows refer only to unconstrained DoFs; values that become zero are
// dropped from the value map.
void process_constraints(System &s) {
  std::set<int32_t> todo;
  for (const auto &r : s.crows) todo.insert(r.first);
  while (!todo.empty())
    for (auto i = todo.begin(); i != todo.end();) {
      ConstraintRow &row = s.crows[*i];
      auto vit = s.cvals.find(*i);
      double rhs = vit == s.cvals.end() ? 0.0 : vit->second;
      std::vector<int32_t> expand;
      for (const auto &it : row)
        if (it.first != *i && s.is_constrained(it.first)) {
          todo.insert(it.first);
          expand.push_back(it.first);
        }
      for (int32_t x : expand) {
        const double c = row[x];
        for (const auto &it : s.crows.at(x)) row[it.first] += it.second * c;
        auto sub = s.cvals.find(x);
        if (sub != s.cvals.end()) rhs += sub->second * c;
        row.erase(x);
      }
      if (rhs != 0.0)
        s.cvals[*i] = rhs;
      else
        s.cvals.erase(*i);
      if (expand.empty())
        i = todo.erase(i);
      else
        ++i;
    }
  s.general = false;
  for (const auto &r : s.crows) s.general |= !r.second.empty();
  for (const auto &v : s.cvals) s.general |= v.second != 0.0;
  s.cons.assign(s.ndofs, 0);
  for (const auto &r : s.crows) s.cons[r.first] = 1;
}

// DofMap::find_connected_dofs (src/base/dof_map.C:2915-2972): add the DoFs
// that constrained DoFs depend on, recursively; the result is sorted.
void find_connected_dofs(const System &s, std::vector<int32_t> &dofs) {
  std::set<int32_t> set(dofs.begin(), dofs.end());
  bool done = true;
  for (int32_t d : dofs) {
    auto pos = s.crows.find(d);
    if (pos == s.crows.end()) continue;
    for (const auto &it : pos->second)
      if (!set.count(it.first)) {
        set.insert(it.first);
        done = false;
      }
  }
  if (!done) {
    dofs.assign(set.begin(), set.end());
    find_connected_dofs(s, dofs);
  }
}

// SparsityPattern::Build with constraints (sparsity_pattern.C:96-147, 150-260):
// each element's DoFs grow by find_connected_dofs, are sorted and made unique,
// and every one of them couples to all of them.
void build_sparsity_general(System &s) {
  std::vector<std::vector<int32_t>> rows(s.ndofs);
  for (i64 e = 0; e < s.m->n_elem; ++e) {
    std::vector<int32_t> d(&s.edofs[(size_t)e * s.dpe], &s.edofs[(size_t)e * s.dpe] + s.dpe);
    find_connected_dofs(s, d);
    std::sort(d.begin(), d.end());
    d.erase(std::unique(d.begin(), d.end()), d.end());
    for (int32_t r : d) rows[r].insert(rows[r].end(), d.begin(), d.end());
  }
  s.rp.assign(s.ndofs + 1, 0);
  for (i64 r = 0; r < s.ndofs; ++r) {
    std::sort(rows[r].begin(), rows[r].end());
    rows[r].erase(std::unique(rows[r].begin(), rows[r].end()), rows[r].end());
    s.rp[r + 1] = s.rp[r] + (int32_t)rows[r].size();
  }
  s.ci.resize(s.rp[s.ndofs]);
  for (i64 r = 0; r < s.ndofs; ++r) std::copy(rows[r].begin(), rows[r].end(), s.ci.begin() + s.rp[r]);
}

// Dense row-major matrix for the constraint algebra (DenseMatrix<Number>).
struct Dense {
  int m = 0, n = 0;
  std::vector<double> a;
  void resize(int mm, int nn) {
    m = mm;
    n = nn;
    a.assign((size_t)m * n, 0.0);
  }
  double &operator()(int i, int j) { return a[(size_t)i * n + j]; }
  double operator()(int i, int j) const { return a[(size_t)i * n + j]; }
};

// A <- A * B (DenseMatrix::right_multiply)
void right_multiply(Dense &A, const Dense &B) {
  Dense R;
  R.resize(A.m, B.n);
  for (int i = 0; i < A.m; ++i)
    for (int j = 0; j < B.n; ++j) {
      double t = 0.0;
      for (int k = 0; k < A.n; ++k) t += A(i, k) * B(k, j);
      R(i, j) = t;
    }
  A = R;
}

// A <- B^T * A (DenseMatrix::left_multiply_transpose)
void left_multiply_transpose(Dense &A, const Dense &B) {
  Dense R;
  R.resize(B.n, A.n);
  for (int i = 0; i < B.n; ++i)
    for (int j = 0; j < A.n; ++j) {
      double t = 0.0;
      for (int k = 0; k < B.m; ++k) t += B(k, i) * A(k, j);
      R(i, j) = t;
    }
  A = R;
}

// DofMap::build_constraint_matrix_and_vector (dof_map_constraints.C:3545-3660):
// x_elem = C x_expanded + H.  The DoFs that constrained DoFs depend on and
// that are not already present are appended (sorted); C(i,i) = 1 for a free
// DoF, C(i, pos(j)) = c_ij and H(i) = value for a constrained one; recurses
// while appended DoFs are constrained themselves.
void build_constraint_matrix_and_vector(const System &s, Dense &C, std::vector<double> &H,
                                        std::vector<int32_t> &dofs, bool recursive) {
  std::set<int32_t> add;
  bool any = false;
  for (int32_t d : dofs) {
    auto pos = s.crows.find(d);
    if (pos == s.crows.end()) continue;
    any = true;
    for (const auto &it : pos->second) add.insert(it.first);
  }
  if (!any) return;
  for (int32_t d : dofs) add.erase(d);
  if (add.empty() && recursive) return;
  const int old = (int)dofs.size();
  dofs.insert(dofs.end(), add.begin(), add.end());
  C.resize(old, (int)dofs.size());
  H.assign(old, 0.0);
  for (int i = 0; i < old; ++i) {
    auto pos = s.crows.find(dofs[i]);
    if (pos == s.crows.end()) {
      C(i, i) = 1.0;
      continue;
    }
    for (const auto &it : pos->second)
      for (int j = 0; j < (int)dofs.size(); ++j)
        if (dofs[j] == it.first) C(i, j) = it.second;
    auto v = s.cvals.find(dofs[i]);
    if (v != s.cvals.end()) H[i] = v->second;
  }
  Dense Cn;
  std::vector<double> Hn;
  build_constraint_matrix_and_vector(s, Cn, Hn, dofs, true);
  if (C.n == Cn.m && Cn.n == (int)dofs.size()) {
    for (int i = 0; i < C.m; ++i)  // H += C Hn
      for (int k = 0; k < C.n; ++k) H[i] += C(i, k) * Hn[k];
    right_multiply(C, Cn);
  }
}

// DofMap::heterogeneously_constrain_element_matrix_and_vector
// (dof_map_constraints.C:2518-2616): F <- C^T (F - K H), K <- C^T K C over the
// expanded DoFs, then each constrained row i is zeroed with K(i,i) = 1 and, with
// asymmetric_constraint_rows, K(i, pos(j)) = -c_ij and F(i) = its value
// (without: F(i) = 0).  K (n x n row-major) and F grow to the expanded size;
// dofs receives the expanded DoF list.
void constrain_general(const System &s, std::vector<int32_t> &dofs, std::vector<double> &K, std::vector<double> &F) {
  const int n = (int)dofs.size();
  Dense C;
  std::vector<double> H;
  build_constraint_matrix_and_vector(s, C, H, dofs, false);
  if (!(C.m == n && C.n == (int)dofs.size())) return;
  Dense Km;
  Km.resize(n, n);
  Km.a = K;
  std::vector<double> G(n);
  for (int i = 0; i < n; ++i) {
    double kh = 0.0;
    for (int j = 0; j < n; ++j) kh += Km(i, j) * H[j];
    G[i] = F[i] - kh;
  }
  const int m = C.n;
  F.assign(m, 0.0);
  for (int j = 0; j < m; ++j) {  // DenseMatrix::vector_mult_transpose
    double t = 0.0;
    for (int i = 0; i < n; ++i) t += C(i, j) * G[i];
    F[j] = t;
  }
  left_multiply_transpose(Km, C);
  right_multiply(Km, C);
  for (int i = 0; i < m; ++i) {
    auto pos = s.crows.find(dofs[i]);
    if (pos == s.crows.end()) continue;
    for (int j = 0; j < m; ++j) Km(i, j) = 0.0;
    Km(i, i) = 1.0;
    if (!s.asymmetric) {
      F[i] = 0.0;
      continue;
    }
    for (const auto &it : pos->second)
      for (int j = 0; j < m; ++j)
        if (dofs[j] == it.first) Km(i, j) = -it.second;
    auto v = s.cvals.find(dofs[i]);
    F[i] = v == s.cvals.end() ? 0.0 : v->second;
  }
  K = Km.a;
}

// EigenSparseMatrix::add_matrix (src/numerics/eigen_sparse_matrix.C:121-135,
// 289-300): coeffRef(i,j) += v per entry in element order; LaspackMatrix::pos
// (laspack_matrix.C:713-732) finds the slot by binary search in the sorted row.
inline i64 slot(const System &s, int32_t r, int32_t c) {
  const int32_t *b = &s.ci[s.rp[r]], *e = &s.ci[s.rp[r + 1]];
  return s.rp[r] + (std::lower_bound(b, e, c) - b);
}

// Poisson element: Ke(i,j) += JxW*kappa*(dphi_i . dphi_j), Fe(i) += JxW*f*phi_i
// (introduction_ex4.C:464-467, 480-537; kappa-weighted as adjoints_ex2/L-shaped.C:101).
bool element_poisson(const System &s, FEValues &fe, i64 e, const double *kappa, int f_kind, double *K,
                     double *F) {
  if (!fe.reinit(*s.m, e)) return false;
  const int n = fe.n;
  const double kap = kappa ? kappa[e] : 1.0;
  std::fill(K, K + n * n, 0.);
  std::fill(F, F + n, 0.);
  for (int qp = 0; qp < fe.nq; ++qp)
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) K[i * n + j] += fe.JxW[qp] * kap * dot3(fe.dp(i, qp), fe.dp(j, qp));
  for (int qp = 0; qp < fe.nq; ++qp) {
    const double f = forcing(f_kind, &fe.xyz[3 * qp]);
    for (int i = 0; i < n; ++i) F[i] += fe.JxW[qp] * f * fe.ph(i, qp);
  }
  return true;
}

// LinearElasticity assembly (systems_of_equations_ex6.C:127-150, 220-236):
// Ke_var[i][k](a,b) += JxW * C_ijkl * dphi_b(l) * dphi_a(j),
// C_ijkl = l1 d_ij d_kl + l2 (d_ik d_jl + d_il d_jk), Fe_var[i](a) += JxW*f_i*phi_a.
bool element_elasticity(const System &s, FEValues &fe, i64 e, double E, double nu, double fz, double *K,
                        double *F) {
  if (!fe.reinit(*s.m, e)) return false;
  const int n = fe.n, N = 3 * n;
  const double lam1 = (E * nu) / ((1. + nu) * (1. - 2. * nu));
  const double lam2 = E / (2. * (1. + nu));
  auto kd = [](int a, int b) { return a == b ? 1.0 : 0.0; };
  double C[3][3][3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l)
          C[i][j][k][l] = lam1 * kd(i, j) * kd(k, l) + lam2 * (kd(i, k) * kd(j, l) + kd(i, l) * kd(j, k));
  std::fill(K, K + N * N, 0.);
  std::fill(F, F + N, 0.);
  const double fv[3] = {0., 0., fz};
  for (int qp = 0; qp < fe.nq; ++qp) {
    for (int a = 0; a < n; ++a)
      for (int b = 0; b < n; ++b)
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k)
              for (int l = 0; l < 3; ++l)
                K[(i * n + a) * N + (k * n + b)] += fe.JxW[qp] * C[i][j][k][l] * fe.dp(b, qp)[l] * fe.dp(a, qp)[j];
    for (int a = 0; a < n; ++a)
      for (int i = 0; i < 3; ++i) F[i * n + a] += fe.JxW[qp] * (fv[i] * fe.ph(a, qp));
  }
  return true;
}

void scatter(const System &s, i64 e, const double *K, const double *F, double *vals, double *rhs) {
  const int n = s.dpe;
  const int32_t *d = &s.edofs[(size_t)e * n];
  for (int i = 0; i < n; ++i) {
    if (vals)
      for (int j = 0; j < n; ++j) vals[slot(s, d[i], d[j])] += K[i * n + j];
    if (rhs) rhs[d[i]] += F[i];  // NumericVector::add_vector (numeric_vector.C:690-697)
  }
}

// mt19937_64 uniform on [0,1): top 53 bits.
inline double unif(std::mt19937_64 &g) { return (double)(g() >> 11) * 0x1.0p-53; }

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
struct orc_mesh {
  Mesh m;
};
struct orc_system {
  System s;
};

extern "C" {

orc_mesh *orc_build_cube(int nx, int ny, int nz, int elem_type) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || nodes_per(elem_type) == 0) return nullptr;
  Mesh *m = build_cube(nx, ny, nz, elem_type);
  orc_mesh *o = new orc_mesh;
  o->m = std::move(*m);
  delete m;
  return o;
}
void orc_mesh_free(orc_mesh *m) { delete m; }
int64_t orc_mesh_n_nodes(const orc_mesh *m) { return m->m.n_nodes; }
int64_t orc_mesh_n_elem(const orc_mesh *m) { return m->m.n_elem; }
int orc_mesh_nodes_per_elem(const orc_mesh *m) { return m->m.npe; }
int orc_mesh_elem_type(const orc_mesh *m) { return m->m.type; }
void orc_mesh_coords(const orc_mesh *m, double *xyz) { std::memcpy(xyz, m->m.xyz.data(), m->m.xyz.size() * 8); }
void orc_mesh_set_coords(orc_mesh *m, const double *xyz) { std::memcpy(m->m.xyz.data(), xyz, m->m.xyz.size() * 8); }
void orc_mesh_conn(const orc_mesh *m, int32_t *conn) { std::memcpy(conn, m->m.conn.data(), m->m.conn.size() * 4); }
void orc_mesh_node_bmask(const orc_mesh *m, uint8_t *mask) { std::memcpy(mask, m->m.bmask.data(), m->m.bmask.size()); }
void orc_mesh_side_bid(const orc_mesh *m, int8_t *bid) { std::memcpy(bid, m->m.side_bid.data(), m->m.side_bid.size()); }

void orc_perturb_interior(orc_mesh *om, uint64_t seed, double frac) {
  Mesh &m = om->m;
  std::mt19937_64 g(seed);
  for (i64 n = 0; n < m.n_nodes; ++n) {
    if (m.bmask[n]) continue;
    for (int c = 0; c < 3; ++c) {
      const double h = 1.0 / (double)m.cells[c];
      m.xyz[3 * n + c] += frac * h * (2.0 * unif(g) - 1.0);
    }
  }
}

void orc_kappa_lognormal(int64_t n_elem, uint64_t seed, double *kappa) {
  std::mt19937_64 g(seed);
  for (i64 e = 0; e < n_elem; ++e) {
    const double u1 = 1.0 - unif(g);  // (0,1]
    const double u2 = unif(g);
    const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
    kappa[e] = std::exp(z);
  }
}

double orc_min_detj(const orc_mesh *om, int npts) {
  const Mesh &m = om->m;
  double mn = 1e300;
  FEValues fe;
  fe.init(m.type, npts);
  for (i64 e = 0; e < m.n_elem; ++e) {
    // detJ at every qp regardless of affine shortcut: evaluate directly
    P3 pt[27];
    for (int a = 0; a < fe.n; ++a) {
      const int32_t nd = m.conn[e * m.npe + a];
      pt[a] = {m.xyz[3 * (size_t)nd], m.xyz[3 * (size_t)nd + 1], m.xyz[3 * (size_t)nd + 2]};
    }
    for (int p = 0; p < fe.nq; ++p) {
      double J[9] = {0};
      for (int a = 0; a < fe.n; ++a) {
        const double *d = &fe.dref[((size_t)a * fe.nq + p) * 3];
        const double X[3] = {pt[a].x, pt[a].y, pt[a].z};
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) J[r * 3 + c] += X[c] * d[r];
      }
      // J[r*3+c] = d x_c / d xi_r; det J by the cofactor expansion of fe_map.C:1046-1048
      const double det2 = J[0] * (J[4] * J[8] - J[5] * J[7]) + J[1] * (J[5] * J[6] - J[3] * J[8]) +
                          J[2] * (J[3] * J[7] - J[4] * J[6]);
      mn = std::min(mn, det2);
    }
  }
  return mn;
}

orc_system *orc_system_create(const orc_mesh *om, int n_vars, int order, unsigned mask) {
  (void)order;  // FE order equals the element's own Lagrange order here
  orc_system *o = new orc_system;
  o->s.m = &om->m;
  o->s.nv = n_vars;
  build_dofs(o->s);
  build_constraints(o->s, mask);
  build_sparsity(o->s);
  return o;
}
void orc_system_free(orc_system *s) { delete s; }
int64_t orc_n_dofs(const orc_system *s) { return s->s.ndofs; }
int64_t orc_nnz(const orc_system *s) { return s->s.rp.empty() ? 0 : s->s.rp.back(); }
int orc_dofs_per_elem(const orc_system *s) { return s->s.dpe; }
void orc_elem_dofs(const orc_system *s, int32_t *d) { std::memcpy(d, s->s.edofs.data(), s->s.edofs.size() * 4); }
void orc_constrained(const orc_system *s, uint8_t *f) { std::memcpy(f, s->s.cons.data(), s->s.cons.size()); }
void orc_pattern(const orc_system *s, int32_t *rp, int32_t *ci) {
  std::memcpy(rp, s->s.rp.data(), s->s.rp.size() * 4);
  std::memcpy(ci, s->s.ci.data(), s->s.ci.size() * 4);
}

int orc_add_constraint_row(orc_system *o, int32_t dof, int n, const int32_t *cols, const double *coefs, double rhs,
                           int forbid) {
  ConstraintRow row;
  for (int k = 0; k < n; ++k) row[cols[k]] = coefs[k];
  return add_constraint_row(o->s, dof, row, rhs, forbid != 0);
}

void orc_set_asymmetric_constraint_rows(orc_system *o, int on) { o->s.asymmetric = on != 0; }

int orc_process_constraints(orc_system *o, int error_on_loop) {
  if (error_on_loop && constraint_loop(o->s)) return -1;
  process_constraints(o->s);
  if (o->s.general)
    build_sparsity_general(o->s);
  else
    build_sparsity(o->s);
  return 0;
}

void orc_constraint_counts(const orc_system *o, int64_t *n_rows, int64_t *n_entries) {
  *n_rows = (int64_t)o->s.crows.size();
  int64_t k = 0;
  for (const auto &r : o->s.crows) k += (int64_t)r.second.size();
  *n_entries = k;
}

void orc_constraint_rows(const orc_system *o, int32_t *dofs, int32_t *rp, int32_t *cols, double *coefs, double *rhs) {
  int64_t r = 0, k = 0;
  rp[0] = 0;
  for (const auto &row : o->s.crows) {
    dofs[r] = row.first;
    auto v = o->s.cvals.find(row.first);
    rhs[r] = v == o->s.cvals.end() ? 0.0 : v->second;
    for (const auto &it : row.second) {
      cols[k] = it.first;
      coefs[k] = it.second;
      ++k;
    }
    rp[++r] = (int32_t)k;
  }
}

int orc_constrain_element(const orc_system *o, int64_t e, const double *K, const double *F, int32_t *dofs_out,
                          double *K_out, double *F_out, int cap) {
  const System &s = o->s;
  const int n = s.dpe;
  std::vector<int32_t> d(&s.edofs[(size_t)e * n], &s.edofs[(size_t)e * n] + n);
  std::vector<double> Kx(K, K + (size_t)n * n), Fx(F, F + n);
  constrain_general(s, d, Kx, Fx);
  const int m = (int)d.size();
  if (m > cap) return -1;
  std::copy(d.begin(), d.end(), dofs_out);
  std::copy(Kx.begin(), Kx.end(), K_out);
  std::copy(Fx.begin(), Fx.end(), F_out);
  return m;
}

void orc_enforce_constraints(const orc_system *o, double *x, int homogeneous) {
  // DofMap::enforce_constraints_exactly (dof_map_constraints.C:3088-3170):
  // x_i = value_i + sum_j c_ij x_j on every constrained DoF, reading the
  // vector as it was on entry
  const std::vector<double> old(x, x + o->s.ndofs);
  for (const auto &row : o->s.crows) {
    double t = 0.0;
    if (!homogeneous) {
      auto v = o->s.cvals.find(row.first);
      if (v != o->s.cvals.end()) t = v->second;
    }
    for (const auto &it : row.second) t += it.second * old[it.first];
    x[row.first] = t;
  }
}

int orc_element_poisson(const orc_system *os, int npts, const double *kappa, int f_kind, int64_t e0,
                        int64_t count, int constrain_flag, double *Ke, double *Fe) {
  const System &s = os->s;
  FEValues fe;
  fe.init(s.m->type, npts);
  const int n = s.dpe;
  for (i64 k = 0; k < count; ++k) {
    double *K = Ke + (size_t)k * n * n, *F = Fe + (size_t)k * n;
    if (!element_poisson(s, fe, e0 + k, kappa, f_kind, K, F)) return -1;
    if (constrain_flag) constrain(s, e0 + k, K, F);
  }
  return 0;
}

// System assembly loop (ImplicitSystem::assemble -> user assembly, introduction_ex4.C:313-560).
int orc_assemble_poisson(const orc_system *os, int npts, const double *kappa, int f_kind, double *vals,
                         double *rhs) {
  const System &s = os->s;
  std::fill(vals, vals + s.rp.back(), 0.);  // matrix->zero(); rhs->zero() (implicit_system.C:94-108)
  std::fill(rhs, rhs + s.ndofs, 0.);
  FEValues fe;
  fe.init(s.m->type, npts);
  const int n = s.dpe;
  std::vector<double> K((size_t)n * n), F(n);
  for (i64 e = 0; e < s.m->n_elem; ++e) {
    if (!element_poisson(s, fe, e, kappa, f_kind, K.data(), F.data())) return -1;
    if (s.general) {
      std::vector<int32_t> d(&s.edofs[(size_t)e * n], &s.edofs[(size_t)e * n] + n);
      std::vector<double> Kx(K), Fx(F);
      constrain_general(s, d, Kx, Fx);
      const int m = (int)d.size();
      for (int i = 0; i < m; ++i) {
        for (int j = 0; j < m; ++j) vals[slot(s, d[i], d[j])] += Kx[(size_t)i * m + j];
        rhs[d[i]] += Fx[i];
      }
      continue;
    }
    constrain(s, e, K.data(), F.data());
    scatter(s, e, K.data(), F.data(), vals, rhs);
  }
  return 0;
}

// FEMSystem::assembly(get_residual) for the Poisson physics
// (fem_system.C:875-1010; element_time_derivative of a Laplace-type physics:
// R_e(i) = sum_qp JxW (kappa grad u . grad phi_i - f phi_i) = (K_e u_e - F_e)(i)),
// then constrain_element_vector's C^T (dof_map_constraints.C:2439-2516: empty
// rows give 0 at constrained DoFs), added in element order.
int orc_residual_poisson(const orc_system *os, int npts, const double *kappa, int f_kind, const double *u,
                         double *R) {
  const System &s = os->s;
  std::fill(R, R + s.ndofs, 0.);
  FEValues fe;
  fe.init(s.m->type, npts);
  const int n = s.dpe;
  std::vector<double> K((size_t)n * n), F(n);
  for (i64 e = 0; e < s.m->n_elem; ++e) {
    if (!element_poisson(s, fe, e, kappa, f_kind, K.data(), F.data())) return -1;
    const int32_t *d = &s.edofs[(size_t)e * n];
    for (int i = 0; i < n; ++i) {
      if (s.cons[d[i]]) continue;
      double r = 0.;
      for (int j = 0; j < n; ++j) r += K[(size_t)i * n + j] * u[d[j]];
      R[d[i]] += r - F[i];
    }
  }
  return 0;
}

// A user physics through FEMSystem::assembly (fem_system.C:875-1010): the
// steady reaction-diffusion -div(kappa grad u) + sigma u^3 = f written as
// FEMSystem physics write element_time_derivative (L-shaped.C:58-108: F and K
// accumulate JxW-weighted terms at each quadrature point, u and grad u from
// FEMContext::interior_value / interior_gradient, fem_context.C:314-380, 411-490):
//   F_e(i)   = sum_qp JxW (kappa_e grad u . grad phi_i + sigma u^3 phi_i - f phi_i)
//   K_e(i,j) = sum_qp JxW (kappa_e grad phi_j . grad phi_i + 3 sigma u^2 phi_j phi_i)
// then the element constraint of the empty Dirichlet rows (constrain above).
// `variant` plants a mistake for the tests that show the pins catch one:
// 1 = Jacobian reaction factor 2 instead of 3, 2 = residual reaction u^2.
static bool element_reaction_diffusion(const System &s, FEValues &fe, i64 e, const double *kappa, double sigma,
                                       int f_kind, const double *u, int variant, double *K, double *F) {
  if (!fe.reinit(*s.m, e)) return false;
  const int n = fe.n;
  const double kap = kappa ? kappa[e] : 1.0;
  const int32_t *d = &s.edofs[(size_t)e * n];
  if (K) std::fill(K, K + n * n, 0.);
  std::fill(F, F + n, 0.);
  for (int qp = 0; qp < fe.nq; ++qp) {
    double uq = 0, g[3] = {0, 0, 0};
    for (int a = 0; a < n; ++a) {
      const double ua = u ? u[d[a]] : 0.0;
      uq += ua * fe.ph(a, qp);
      const double *dp = fe.dp(a, qp);
      g[0] += ua * dp[0];
      g[1] += ua * dp[1];
      g[2] += ua * dp[2];
    }
    const double w = fe.JxW[qp], f = forcing(f_kind, &fe.xyz[3 * qp]);
    const double react = variant == 2 ? sigma * uq * uq : sigma * uq * uq * uq;
    const double dreact = (variant == 1 ? 2.0 : 3.0) * sigma * uq * uq;
    for (int i = 0; i < n; ++i) {
      const double pi = fe.ph(i, qp);
      F[i] += w * (kap * dot3(g, fe.dp(i, qp)) + react * pi - f * pi);
      if (K)
        for (int j = 0; j < n; ++j)
          K[i * n + j] += w * (kap * dot3(fe.dp(j, qp), fe.dp(i, qp)) + dreact * fe.ph(j, qp) * pi);
    }
  }
  return true;
}

// Element results (constrained when constrain_flag is set) of elements
// e0 .. e0+count-1; Ke may be NULL (residual only).
int orc_fem_element_reaction_diffusion(const orc_system *os, int npts, const double *kappa, double sigma,
                                       int f_kind, const double *u, int64_t e0, int64_t count, int constrain_flag,
                                       int variant, double *Ke, double *Fe) {
  const System &s = os->s;
  if (s.nv != 1 || s.general) return -2;
  FEValues fe;
  fe.init(s.m->type, npts);
  const int n = s.dpe;
  for (i64 k = 0; k < count; ++k) {
    double *K = Ke ? Ke + (size_t)k * n * n : nullptr, *F = Fe + (size_t)k * n;
    if (!element_reaction_diffusion(s, fe, e0 + k, kappa, sigma, f_kind, u, variant, K, F)) return -1;
    if (constrain_flag) {
      std::vector<double> Kt((size_t)n * n);
      constrain(s, e0 + k, K ? K : Kt.data(), F);
    }
  }
  return 0;
}

// FEMSystem::assembly(get_residual = R != NULL, get_jacobian = vals != NULL)
// for that physics: constrained element results summed in element order into
// the CSR values and the residual (both zeroed first).
int orc_fem_reaction_diffusion(const orc_system *os, int npts, const double *kappa, double sigma, int f_kind,
                               const double *u, int variant, double *vals, double *R) {
  const System &s = os->s;
  if (s.nv != 1 || s.general) return -2;
  if (vals) std::fill(vals, vals + s.rp.back(), 0.);
  if (R) std::fill(R, R + s.ndofs, 0.);
  FEValues fe;
  fe.init(s.m->type, npts);
  const int n = s.dpe;
  std::vector<double> K((size_t)n * n), F(n);
  for (i64 e = 0; e < s.m->n_elem; ++e) {
    if (!element_reaction_diffusion(s, fe, e, kappa, sigma, f_kind, u, variant, K.data(), F.data())) return -1;
    constrain(s, e, K.data(), F.data());
    scatter(s, e, K.data(), F.data(), vals, R);
  }
  return 0;
}

int orc_element_elasticity(const orc_system *os, int npts, double E, double nu, double fz, int64_t e0,
                           int64_t count, int constrain_flag, double *Ke, double *Fe) {
  const System &s = os->s;
  FEValues fe;
  fe.init(s.m->type, npts);
  const int n = s.dpe;
  for (i64 k = 0; k < count; ++k) {
    double *K = Ke + (size_t)k * n * n, *F = Fe + (size_t)k * n;
    if (!element_elasticity(s, fe, e0 + k, E, nu, fz, K, F)) return -1;
    if (constrain_flag) constrain(s, e0 + k, K, F);
  }
  return 0;
}

int orc_assemble_elasticity(const orc_system *os, int npts, double E, double nu, double fz, double *vals,
                            double *rhs) {
  const System &s = os->s;
  std::fill(vals, vals + s.rp.back(), 0.);
  std::fill(rhs, rhs + s.ndofs, 0.);
  FEValues fe;
  fe.init(s.m->type, npts);
  const int n = s.dpe;
  std::vector<double> K((size_t)n * n), F(n);
  for (i64 e = 0; e < s.m->n_elem; ++e) {
    if (!element_elasticity(s, fe, e, E, nu, fz, K.data(), F.data())) return -1;
    constrain(s, e, K.data(), F.data());
    scatter(s, e, K.data(), F.data(), vals, rhs);
  }
  return 0;
}

void orc_spmv(int64_t n, const int32_t *rp, const int32_t *ci, const double *v, const double *x, double *y) {
#pragma omp parallel for schedule(static)
  for (i64 r = 0; r < n; ++r) {
    double t = 0;
    for (i64 k = rp[r]; k < rp[r + 1]; ++k) t += v[k] * x[ci[k]];
    y[r] = t;
  }
}

static double dotv(const std::vector<double> &a, const std::vector<double> &b) {
  double t = 0;
  for (size_t i = 0; i < a.size(); ++i) t += a[i] * b[i];
  return t;
}

// Eigen::internal::conjugate_gradient (ConjugateGradient.h:28-90) with
// DiagonalPreconditioner (BasicPreconditioners.h:66-76: 1/a_jj, or 1 when the
// diagonal is absent or zero).  Stops when |r|^2 < max(tol^2 |b|^2, DBL_MIN);
// reports iterations and sqrt(|r|^2/|b|^2).
// precond: 1 JACOBI_PRECOND (DiagonalPreconditioner), 0 IDENTITY_PRECOND
// (IdentityPreconditioner, BasicPreconditioners.h:200-226: z = r), the two
// EigenSparseLinearSolver::solve offers without a factorisation
// (eigen_sparse_linear_solver.C:136-170, 177-205).
static std::vector<double> precond_diag(int64_t n, const int32_t *rp, const int32_t *ci, const double *v,
                                        int precond) {
  std::vector<double> invd(n, 1.0);
  if (precond == 1)
    for (i64 r = 0; r < n; ++r)
      for (i64 k = rp[r]; k < rp[r + 1]; ++k)
        if (ci[k] == r && v[k] != 0.0) invd[r] = 1.0 / v[k];
  return invd;
}

void orc_cg_eigen_pc(int64_t n, const int32_t *rp, const int32_t *ci, const double *v, const double *b, double *x,
                     double tol, int64_t maxits, int precond, int64_t *iters, double *rel_err) {
  const std::vector<double> invd = precond_diag(n, rp, ci, v, precond);
  std::vector<double> bb(b, b + n), xx(x, x + n), res(n), tmp(n), p(n), z(n);
  orc_spmv(n, rp, ci, v, x, tmp.data());
  for (i64 i = 0; i < n; ++i) res[i] = bb[i] - tmp[i];
  const double rhsNorm2 = dotv(bb, bb);
  if (rhsNorm2 == 0) {
    std::fill(x, x + n, 0.);
    *iters = 0;
    *rel_err = 0;
    return;
  }
  const double threshold = std::max(tol * tol * rhsNorm2, 2.2250738585072014e-308);
  double residualNorm2 = dotv(res, res);
  if (residualNorm2 < threshold) {
    *iters = 0;
    *rel_err = std::sqrt(residualNorm2 / rhsNorm2);
    return;
  }
  for (i64 i = 0; i < n; ++i) p[i] = invd[i] * res[i];
  double absNew = dotv(res, p);
  i64 it = 0;
  while (it < maxits) {
    orc_spmv(n, rp, ci, v, p.data(), tmp.data());
    const double alpha = absNew / dotv(p, tmp);
    for (i64 i = 0; i < n; ++i) xx[i] += alpha * p[i];
    for (i64 i = 0; i < n; ++i) res[i] -= alpha * tmp[i];
    residualNorm2 = dotv(res, res);
    if (residualNorm2 < threshold) break;
    for (i64 i = 0; i < n; ++i) z[i] = invd[i] * res[i];
    const double absOld = absNew;
    absNew = dotv(res, z);
    const double beta = absNew / absOld;
    for (i64 i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    it++;
  }
  std::memcpy(x, xx.data(), n * 8);
  *iters = it;
  *rel_err = std::sqrt(residualNorm2 / rhsNorm2);
}

// Eigen::BiCGSTAB with DiagonalPreconditioner as EigenSparseLinearSolver::solve
// runs it for BICGSTAB + JACOBI_PRECOND (eigen_sparse_linear_solver.C:175-205;
// Eigen 3.3.9 IterativeLinearSolvers/BiCGSTAB.h:28-112): restart with r0 = r when
// |r0.r| < eps^2 |r0|^2 (the first restart also resets the iteration count);
// stops when |r|^2 <= tol^2 |b|^2.  Returns the iteration count; rel_err =
// |r|/|b|.
void orc_bicgstab_eigen_pc(int64_t n, const int32_t *rp, const int32_t *ci, const double *v, const double *b,
                           double *x, double tol, int64_t maxits, int precond, int64_t *iters, double *rel_err) {
  const std::vector<double> invd = precond_diag(n, rp, ci, v, precond);
  std::vector<double> bb(b, b + n), xx(x, x + n), r(n), r0(n), vv(n, 0.), p(n, 0.), y(n), z(n), s(n), t(n), tmp(n);
  orc_spmv(n, rp, ci, v, x, tmp.data());
  for (i64 i = 0; i < n; ++i) r[i] = bb[i] - tmp[i];
  r0 = r;
  double r0_sqnorm = dotv(r0, r0);
  const double rhs_sqnorm = dotv(bb, bb);
  if (rhs_sqnorm == 0) {
    std::fill(x, x + n, 0.);
    *iters = 0;
    *rel_err = 0;
    return;
  }
  double rho = 1, alpha = 1, w = 1;
  const double tol2 = tol * tol * rhs_sqnorm;
  const double eps = 2.220446049250313e-16, eps2 = eps * eps;
  i64 i = 0, restarts = 0;
  while (dotv(r, r) > tol2 && i < maxits) {
    const double rho_old = rho;
    rho = dotv(r0, r);
    if (std::fabs(rho) < eps2 * r0_sqnorm) {
      orc_spmv(n, rp, ci, v, xx.data(), tmp.data());
      for (i64 k = 0; k < n; ++k) r[k] = bb[k] - tmp[k];
      r0 = r;
      rho = r0_sqnorm = dotv(r, r);
      if (restarts++ == 0) i = 0;
    }
    const double beta = (rho / rho_old) * (alpha / w);
    for (i64 k = 0; k < n; ++k) p[k] = r[k] + beta * (p[k] - w * vv[k]);
    for (i64 k = 0; k < n; ++k) y[k] = invd[k] * p[k];
    orc_spmv(n, rp, ci, v, y.data(), vv.data());
    alpha = rho / dotv(r0, vv);
    for (i64 k = 0; k < n; ++k) s[k] = r[k] - alpha * vv[k];
    for (i64 k = 0; k < n; ++k) z[k] = invd[k] * s[k];
    orc_spmv(n, rp, ci, v, z.data(), t.data());
    const double tmpn = dotv(t, t);
    w = tmpn > 0 ? dotv(t, s) / tmpn : 0.0;
    for (i64 k = 0; k < n; ++k) xx[k] += alpha * y[k] + w * z[k];
    for (i64 k = 0; k < n; ++k) r[k] = s[k] - w * t[k];
    ++i;
  }
  std::memcpy(x, xx.data(), n * 8);
  *iters = i;
  *rel_err = std::sqrt(dotv(r, r) / rhs_sqnorm);
}

void orc_cg_eigen(int64_t n, const int32_t *rp, const int32_t *ci, const double *v, const double *b, double *x,
                  double tol, int64_t maxits, int64_t *iters, double *rel_err) {
  orc_cg_eigen_pc(n, rp, ci, v, b, x, tol, maxits, 1, iters, rel_err);
}

void orc_bicgstab_eigen(int64_t n, const int32_t *rp, const int32_t *ci, const double *v, const double *b, double *x,
                        double tol, int64_t maxits, int64_t *iters, double *rel_err) {
  orc_bicgstab_eigen_pc(n, rp, ci, v, b, x, tol, maxits, 1, iters, rel_err);
}

// Laspack CGIter with JacobiPrecond, Omega = 1 (itersolv.C:377-476, precond.C:26-46):
// RTCResult stops when |r| < eps |b| (rtc.C:48-72), checked before each iteration.
// LaspackLinearSolver::solve zeroes the initial guess (laspack_linear_solver.C:130).
void orc_cg_laspack(int64_t n, const int32_t *rp, const int32_t *ci, const double *v, const double *b, double *x,
                    double tol, int64_t maxits, int64_t *iters, double *rel_res) {
  std::vector<double> d(n, 0.0);
  for (i64 r = 0; r < n; ++r)
    for (i64 k = rp[r]; k < rp[r + 1]; ++k)
      if (ci[k] == r) d[r] = v[k];
  std::vector<double> bb(b, b + n), xx(n, 0.0), r(bb), p(n), q(n), z(n);
  const double bNorm = std::sqrt(dotv(bb, bb));
  double rho = 0, rhoOld = 0;
  i64 it = 0;
  double rn = std::sqrt(dotv(r, r));
  while (!(rn < tol * bNorm) && it < maxits) {
    it++;
    for (i64 i = 0; i < n; ++i) z[i] = 1.0 * (r[i] / d[i]);
    rho = dotv(r, z);
    if (it == 1) {
      p = z;
    } else {
      const double beta = rho / rhoOld;
      for (i64 i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
    orc_spmv(n, rp, ci, v, p.data(), q.data());
    const double alpha = rho / dotv(p, q);
    for (i64 i = 0; i < n; ++i) xx[i] += alpha * p[i];
    for (i64 i = 0; i < n; ++i) r[i] -= alpha * q[i];
    rhoOld = rho;
    rn = std::sqrt(dotv(r, r));
  }
  std::memcpy(x, xx.data(), n * 8);
  *iters = it;
  *rel_res = bNorm != 0 ? rn / bNorm : 1.0;
}

// QoI J = int_{[lo,hi]^3} u dx with a per-qp inclusion test on physical xyz
// (adjoints_ex2/L-qoi.C element_qoi / element_qoi_derivative pattern);
// System::assemble_qoi_derivative with apply_constraints zeroes constrained
// entries (enforce_adjoint_constraints_exactly, dof_map_constraints.C:3261).
static bool inbox(const double *x, double lo, double hi) {
  return x[0] >= lo && x[0] <= hi && x[1] >= lo && x[1] <= hi && x[2] >= lo && x[2] <= hi;
}

int orc_qoi_box_derivative(const orc_system *os, int npts, double lo, double hi, double *dq) {
  const System &s = os->s;
  std::fill(dq, dq + s.ndofs, 0.);
  FEValues fe;
  fe.init(s.m->type, npts);
  std::vector<double> Q(s.dpe);
  for (i64 e = 0; e < s.m->n_elem; ++e) {
    if (!fe.reinit(*s.m, e)) return -1;
    std::fill(Q.begin(), Q.end(), 0.);
    for (int qp = 0; qp < fe.nq; ++qp)
      if (inbox(&fe.xyz[3 * qp], lo, hi))
        for (int i = 0; i < fe.n; ++i) Q[i] += fe.JxW[qp] * fe.ph(i, qp);
    for (int i = 0; i < s.dpe; ++i) dq[s.edofs[(size_t)e * s.dpe + i]] += Q[i];
  }
  for (i64 i = 0; i < s.ndofs; ++i)
    if (s.cons[i]) dq[i] = 0.;
  return 0;
}

double orc_qoi_box(const orc_system *os, int npts, double lo, double hi, const double *u) {
  const System &s = os->s;
  FEValues fe;
  fe.init(s.m->type, npts);
  double J = 0;
  for (i64 e = 0; e < s.m->n_elem; ++e) {
    if (!fe.reinit(*s.m, e)) return NAN;
    for (int qp = 0; qp < fe.nq; ++qp)
      if (inbox(&fe.xyz[3 * qp], lo, hi)) {
        double uq = 0;
        for (int i = 0; i < fe.n; ++i) uq += fe.ph(i, qp) * u[s.edofs[(size_t)e * s.dpe + i]];
        J += fe.JxW[qp] * uq;
      }
  }
  return J;
}

// ImplicitSystem::adjoint_qoi_parameter_sensitivity (implicit_system.C:517-623)
// with assemble_residual_derivatives (:474-512): for each parameter p_j,
// delta = TOLERANCE * max(|p_j|, 1e-3) with TOLERANCE = 1e-6
// (libmesh_common.h:172); s_j = (R(p-delta) - R(p+delta)) / (2 delta) with
// R = A u - b (linear_implicit_system.C:367-381); sens_j = dq/dp_j + s_j . z.
// dq/dkappa = 0 here (J depends on u only).  Only element e's contribution to
// A depends on kappa_e and b does not depend on kappa, so R(p-d) - R(p+d)
// restricted to element e's DoFs is Kc_e(k-d) u_e - Kc_e(k+d) u_e, where Kc_e is
// the constrained element matrix.
int orc_adjoint_sensitivity(const orc_system *os, int npts, const double *kappa, int f_kind, const double *u,
                            const double *z, const int64_t *elems, int64_t n_elems, double *sens) {
  const System &s = os->s;
  FEValues fe;
  fe.init(s.m->type, npts);
  const int n = s.dpe;
  std::vector<double> Km((size_t)n * n), Kp((size_t)n * n), F(n);
  std::vector<double> kap(kappa, kappa + s.m->n_elem);
  for (i64 k = 0; k < n_elems; ++k) {
    const i64 e = elems[k];
    const double p = kappa[e];
    const double delta = 1e-6 * std::max(std::abs(p), 1e-3);
    kap[e] = p - delta;
    if (!element_poisson(s, fe, e, kap.data(), f_kind, Km.data(), F.data())) return -1;
    constrain(s, e, Km.data(), F.data());
    kap[e] = p + delta;
    if (!element_poisson(s, fe, e, kap.data(), f_kind, Kp.data(), F.data())) return -1;
    constrain(s, e, Kp.data(), F.data());
    kap[e] = p;
    const int32_t *d = &s.edofs[(size_t)e * n];
    double acc = 0;
    for (int i = 0; i < n; ++i) {
      double rm = 0, rpv = 0;
      for (int j = 0; j < n; ++j) {
        rm += Km[i * n + j] * u[d[j]];
        rpv += Kp[i * n + j] * u[d[j]];
      }
      acc += ((rm - rpv) / (2 * delta)) * z[d[i]];
    }
    sens[k] = 0.0 + acc;
  }
  return 0;
}

int orc_adjoint_exact(const orc_system *os, int npts, const double *kappa, const double *u, const double *z,
                      double *sens) {
  const System &s = os->s;
  FEValues fe;
  fe.init(s.m->type, npts);
  const int n = s.dpe;
  std::vector<double> K((size_t)n * n), F(n);
  for (i64 e = 0; e < s.m->n_elem; ++e) {
    if (!element_poisson(s, fe, e, kappa, 0, K.data(), F.data())) return -1;
    constrain(s, e, K.data(), F.data());
    const int32_t *d = &s.edofs[(size_t)e * n];
    double acc = 0;
    for (int i = 0; i < n; ++i) {
      if (s.cons[d[i]]) continue;  // constrained rows do not depend on kappa
      double t = 0;
      for (int j = 0; j < n; ++j) t += K[i * n + j] * u[d[j]];
      acc += z[d[i]] * t;
    }
    sens[e] = -acc / kappa[e];
  }
  return 0;
}

void orc_qgauss_1d(int npts, double *x, double *w) { qgauss_1d(npts, x, w); }
void orc_qgauss_hex(int npts, double *xyz, double *w) {
  QRule r = qgauss_hex(npts);
  std::memcpy(xyz, r.p.data(), r.p.size() * 8);
  std::memcpy(w, r.w.data(), r.w.size() * 8);
}
int orc_qgauss_tet(int order, double *xyz, double *w) {
  QRule r = qgauss_tet(order);
  if (xyz) std::memcpy(xyz, r.p.data(), r.p.size() * 8);
  if (w) std::memcpy(w, r.w.data(), r.w.size() * 8);
  return r.n;
}
int orc_shapes(int elem_type, int order, int nq, const double *pts, double *phi, double *dphi) {
  (void)order;
  std::vector<double> a, b;
  const int n = shapes(elem_type, nq, pts, a, b);
  std::memcpy(phi, a.data(), a.size() * 8);
  std::memcpy(dphi, b.data(), b.size() * 8);
  return n;
}

}  // extern "C"
