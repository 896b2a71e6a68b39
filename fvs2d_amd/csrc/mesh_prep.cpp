// Host-side mesh preparation (see mesh_prep.hpp).
//
// The numbering rules are the reference's (src/grid_procs.f90:401-791):
//   * edges are created walking cells in order and each cell's local edges
//     e = (v_e, v_e+1) in order; an edge is created by the lower-numbered cell
//     (or by the only cell on a boundary), so c1 < c2, n1->n2 follows c1's
//     counter-clockwise order and n = (dy,-dx)/|d| points out of c1;
//   * interior edges are processed in ascending edge id, boundary edges per
//     boundary in the order of the listed cells and their local edges.
// The algorithms are this library's own: neighbours come from a hash of
// directed edges instead of the reference's search over node->cell lists,
// node-neighbour stencils from sort+unique, k-nearest centres from a bucket
// grid instead of a k-d tree.
#include "mesh_prep.hpp"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <exception>
#include <thread>
#include <type_traits>
#include <unordered_map>

namespace fvs {
namespace {

constexpr double kPi = 3.14159265358979323846;

std::string fmt(const char *f, long a = 0, long b = 0, long c = 0) {
  char buf[256];
  snprintf(buf, sizeof buf, f, a, b, c);
  return buf;
}

inline double tri_area(double x1, double x2, double x3, double y1, double y2, double y3) {
  return 0.5 * (x1 * (y2 - y3) + x2 * (y3 - y1) + x3 * (y1 - y2));
}

// position of canonical node v in cell c's own frame
inline bool node_in_cell(const MeshPrep &m, int c, int v, double &x, double &y) {
  for (int k = 0; k < m.nvrt[c]; ++k)
    if (m.cnodec[4 * c + k] == v) {
      int p = m.cnode[4 * c + k];
      x = m.nxy[2 * p];
      y = m.nxy[2 * p + 1];
      return true;
    }
  return false;
}

// shift carrying cell j's centroid into cell i's frame (0 unless periodic)
inline void pair_shift(const MeshPrep &m, int i, int j, double &sx, double &sy) {
  sx = 0.0;
  sy = 0.0;
  if (!m.periodic) return;
  for (int a = 0; a < m.nvrt[i]; ++a)
    for (int b = 0; b < m.nvrt[j]; ++b)
      if (m.cnodec[4 * i + a] == m.cnodec[4 * j + b]) {
        int pi = m.cnode[4 * i + a], pj = m.cnode[4 * j + b];
        sx = m.nxy[2 * pi] - m.nxy[2 * pj];
        sy = m.nxy[2 * pi + 1] - m.nxy[2 * pj + 1];
        return;
      }
}

std::string topology(MeshPrep &m) {
  const int N = m.ncells;
  // centroids (vertex averages) and signed areas: src/grid_procs.f90:185-227
  m.cx.assign(N, 0.0);
  m.cy.assign(N, 0.0);
  m.vol.assign(N, 0.0);
  for (int c = 0; c < N; ++c) {
    const int *nd = &m.cnode[4 * c];
    double sx = 0.0, sy = 0.0;
    for (int k = 0; k < m.nvrt[c]; ++k) {
      sx = sx + m.nxy[2 * nd[k]];
      sy = sy + m.nxy[2 * nd[k] + 1];
    }
    double d = m.nvrt[c] == 3 ? 3.0 : 4.0;
    m.cx[c] = sx / d;
    m.cy[c] = sy / d;
    double x1 = m.nxy[2 * nd[0]], y1 = m.nxy[2 * nd[0] + 1], x2 = m.nxy[2 * nd[1]], y2 = m.nxy[2 * nd[1] + 1];
    double x3 = m.nxy[2 * nd[2]], y3 = m.nxy[2 * nd[2] + 1];
    if (m.nvrt[c] == 3) {
      m.vol[c] = tri_area(x1, x2, x3, y1, y2, y3);
    } else {
      double x4 = m.nxy[2 * nd[3]], y4 = m.nxy[2 * nd[3] + 1];
      m.vol[c] = tri_area(x1, x2, x3, y1, y2, y3) + tri_area(x1, x3, x4, y1, y3, y4);
    }
    if (!(m.vol[c] > 0.0)) return fmt("cell %ld is not counter-clockwise (area %ld)", c, 0);
  }
  // node -> cells over canonical ids, cells ascending
  m.n2c_ptr.assign(m.nnodes + 1, 0);
  for (int c = 0; c < N; ++c)
    for (int k = 0; k < m.nvrt[c]; ++k) m.n2c_ptr[m.cnodec[4 * c + k] + 1]++;
  for (int v = 0; v < m.nnodes; ++v) m.n2c_ptr[v + 1] += m.n2c_ptr[v];
  m.n2c.assign(m.n2c_ptr[m.nnodes], 0);
  {
    std::vector<int> fill(m.n2c_ptr.begin(), m.n2c_ptr.end() - 1);
    for (int c = 0; c < N; ++c)
      for (int k = 0; k < m.nvrt[c]; ++k) m.n2c[fill[m.cnodec[4 * c + k]]++] = c;
  }
  // neighbours across local edges: every directed edge (a, b) of every
  // cell, sorted by its node pair; the neighbour across (a, b) owns (b, a)
  auto key = [](int a, int b) { return (uint64_t(uint32_t(a)) << 32) | uint32_t(b); };
  std::vector<std::pair<uint64_t, int64_t>> dir;
  dir.reserve(size_t(N) * 4);
  for (int c = 0; c < N; ++c) {
    int nv = m.nvrt[c];
    for (int e = 0; e < nv; ++e)
      dir.emplace_back(key(m.cnodec[4 * c + e], m.cnodec[4 * c + (e + 1) % nv]), int64_t(c) * 4 + e);
  }
  std::sort(dir.begin(), dir.end());
  for (size_t q = 1; q < dir.size(); ++q)
    if (dir[q].first == dir[q - 1].first)
      return fmt("directed edge (%ld,%ld) appears twice: mesh is not a consistently oriented manifold",
                 long(dir[q].first >> 32), long(dir[q].first & 0xffffffffu));
  m.nghbre.assign(4 * N, -1);
  std::vector<int> nb_local(4 * N, -1);  // local edge index in the neighbour
  for (int c = 0; c < N; ++c) {
    int nv = m.nvrt[c];
    for (int e = 0; e < nv; ++e) {
      const uint64_t rk = key(m.cnodec[4 * c + (e + 1) % nv], m.cnodec[4 * c + e]);
      auto it = std::lower_bound(dir.begin(), dir.end(), std::make_pair(rk, int64_t(-1)));
      if (it != dir.end() && it->first == rk) {
        m.nghbre[4 * c + e] = int(it->second / 4);
        nb_local[4 * c + e] = int(it->second % 4);
      }
    }
  }
  std::vector<std::pair<uint64_t, int64_t>>().swap(dir);
  // edge numbering: src/grid_procs.f90:401-624
  int ne = 0;
  m.cedge.assign(4 * N, -1);
  for (int c = 0; c < N; ++c)
    for (int e = 0; e < m.nvrt[c]; ++e) {
      int nb = m.nghbre[4 * c + e];
      if (nb > c || nb == -1) ++ne;
    }
  m.nedges = ne;
  m.en1.resize(ne); m.en2.resize(ne); m.ec1.resize(ne); m.ec2.resize(ne);
  ne = 0;
  for (int c = 0; c < N; ++c) {
    int nv = m.nvrt[c];
    for (int e = 0; e < nv; ++e) {
      int nb = m.nghbre[4 * c + e];
      if (nb > c || nb == -1) {
        m.en1[ne] = m.cnode[4 * c + e];
        m.en2[ne] = m.cnode[4 * c + (e + 1) % nv];
        m.ec1[ne] = c;
        m.ec2[ne] = nb;
        m.cedge[4 * c + e] = ne++;
      } else {
        m.cedge[4 * c + e] = m.cedge[4 * nb + nb_local[4 * c + e]];
      }
    }
  }
  // edge geometry: src/grid_procs.f90:630-647
  m.ex.resize(ne); m.ey.resize(ne); m.ea.resize(ne); m.enx.resize(ne); m.eny.resize(ne);
  m.esx.assign(ne, 0.0); m.esy.assign(ne, 0.0);
  for (int i = 0; i < ne; ++i) {
    int v1 = m.en1[i], v2 = m.en2[i];
    double dx = m.nxy[2 * v2] - m.nxy[2 * v1], dy = m.nxy[2 * v2 + 1] - m.nxy[2 * v1 + 1];
    m.ea[i] = std::sqrt(dx * dx + dy * dy);
    m.ex[i] = 0.5 * (m.nxy[2 * v1] + m.nxy[2 * v2]);
    m.ey[i] = 0.5 * (m.nxy[2 * v1 + 1] + m.nxy[2 * v2 + 1]);
    m.enx[i] = dy / m.ea[i];
    m.eny[i] = -dx / m.ea[i];
    if (m.periodic && m.ec2[i] >= 0) {
      double px, py;
      node_in_cell(m, m.ec2[i], m.canon[v1], px, py);
      m.esx[i] = m.nxy[2 * v1] - px;
      m.esy[i] = m.nxy[2 * v1 + 1] - py;
    }
  }
  m.nrmlsign.assign(4 * N, 1.0);
  for (int c = 0; c < N; ++c)
    for (int e = 0; e < m.nvrt[c]; ++e)
      if (m.ec1[m.cedge[4 * c + e]] != c) m.nrmlsign[4 * c + e] = -1.0;
  // interior / boundary cells and edges: src/grid_procs.f90:697-791
  m.cell_intr.clear();
  m.ncells_bndr = 0;
  for (int c = 0; c < N; ++c) {
    bool inner = true;
    for (int e = 0; e < m.nvrt[c]; ++e) inner = inner && m.nghbre[4 * c + e] != -1;
    if (inner) m.cell_intr.push_back(c);
    else ++m.ncells_bndr;
  }
  long listed = 0;
  for (int b = 0; b < m.nb; ++b) listed += m.bnd_ncells[b];
  if (listed != m.ncells_bndr)
    return fmt("#s of boundary cells determined %ld != listed in the .bc data %ld", m.ncells_bndr, listed);
  m.edge_intr.clear();
  m.nedges_bndr = 0;
  for (int i = 0; i < ne; ++i) {
    if (m.ec2[i] >= 0) m.edge_intr.push_back(i);
    else ++m.nedges_bndr;
  }
  m.bnd_nedges.assign(m.nb, 0);
  m.bnd_eptr.assign(m.nb + 1, 0);
  m.bnd_edges.clear();
  for (int b = 0; b < m.nb; ++b) {
    m.bnd_eptr[b] = int(m.bnd_edges.size());
    for (int i = 0; i < m.bnd_ncells[b]; ++i) {
      int c = m.bnd_cells[m.bnd_cptr[b] + i];
      if (c < 0 || c >= N) return fmt("boundary %ld lists cell %ld outside the mesh", b + 1, c);
      for (int e = 0; e < m.nvrt[c]; ++e) {
        int je = m.cedge[4 * c + e];
        if (m.ec1[je] == c && m.ec2[je] == -1) {
          m.bnd_edges.push_back(je);
          m.bnd_nedges[b]++;
        }
      }
    }
  }
  m.bnd_eptr[m.nb] = int(m.bnd_edges.size());
  if (long(m.bnd_edges.size()) != m.nedges_bndr)
    return fmt("#s of boundary edges determined %ld != found through the .bc cells %ld", m.nedges_bndr,
               long(m.bnd_edges.size()));
  // src/residual.f90:115-125 pairs bndry%cell(i) with bndry%edge(i)
  for (int b = 0; b < m.nb; ++b)
    if (m.bnd_nedges[b] > m.bnd_ncells[b])
      return fmt("boundary %ld has %ld boundary edges for %ld listed cells; the cell/edge pairing needs one "
                 "boundary edge per listed cell",
                 b + 1, m.bnd_nedges[b], m.bnd_ncells[b]);
  return "";
}

// ---- k nearest cell centres: uniform bucket grid, ties by ascending id ----
struct Buckets {
  double x0, y0, h;
  int nx, ny;
  std::vector<int> ptr, ids;
  void build(const std::vector<double> &x, const std::vector<double> &y) {
    int n = int(x.size());
    double xa = *std::min_element(x.begin(), x.end()), xb = *std::max_element(x.begin(), x.end());
    double ya = *std::min_element(y.begin(), y.end()), yb = *std::max_element(y.begin(), y.end());
    double span = std::max(xb - xa, yb - ya);
    int side = std::max(1, int(std::sqrt(double(n) / 2.0)));
    h = span > 0 ? span / side : 1.0;
    x0 = xa; y0 = ya;
    nx = std::max(1, int((xb - xa) / h) + 1);
    ny = std::max(1, int((yb - ya) / h) + 1);
    ptr.assign(size_t(nx) * ny + 1, 0);
    std::vector<int> cell_of(n);
    for (int i = 0; i < n; ++i) {
      int bx = std::min(nx - 1, int((x[i] - x0) / h)), by = std::min(ny - 1, int((y[i] - y0) / h));
      cell_of[i] = by * nx + bx;
      ptr[cell_of[i] + 1]++;
    }
    for (size_t k = 0; k + 1 < ptr.size(); ++k) ptr[k + 1] += ptr[k];
    ids.assign(n, 0);
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int i = 0; i < n; ++i) ids[fill[cell_of[i]]++] = i;
  }
};

void knearest(const Buckets &B, const std::vector<double> &x, const std::vector<double> &y, double qx, double qy,
              int k, std::vector<std::pair<double, int>> &out) {
  out.clear();
  int bx = std::min(B.nx - 1, std::max(0, int((qx - B.x0) / B.h)));
  int by = std::min(B.ny - 1, std::max(0, int((qy - B.y0) / B.h)));
  for (int ring = 0;; ++ring) {
    for (int j = by - ring; j <= by + ring; ++j)
      for (int i = bx - ring; i <= bx + ring; ++i) {
        if (i < 0 || j < 0 || i >= B.nx || j >= B.ny) continue;
        if (std::max(std::abs(i - bx), std::abs(j - by)) != ring) continue;
        int b = j * B.nx + i;
        for (int p = B.ptr[b]; p < B.ptr[b + 1]; ++p) {
          int id = B.ids[p];
          double dx = x[id] - qx, dy = y[id] - qy;
          out.push_back({dx * dx + dy * dy, id});
        }
      }
    std::sort(out.begin(), out.end());
    if (out.size() > size_t(k)) out.resize(k);
    // every point outside the searched square is farther than ring*h
    double reach = ring * B.h;
    bool covered = (bx - ring <= 0 && by - ring <= 0 && bx + ring >= B.nx - 1 && by + ring >= B.ny - 1);
    if ((out.size() == size_t(k) && out.back().first < reach * reach) || covered) return;
  }
}

std::string lsq_stencil(MeshPrep &m, bool nn, double p) {
  const int N = m.ncells;
  m.lsq_ptr.assign(N + 1, 0);
  if (!nn) {
    // setup_fn (src/gradient_lsq.f90:88-131): the vertex-opposite neighbour
    // order nghbr(1..nvrt), where nghbr(in) lies across local edge in-2
    for (int c = 0; c < N; ++c) m.lsq_ptr[c + 1] = m.lsq_ptr[c] + m.nvrt[c];
    m.lsq_cell.assign(m.lsq_ptr[N], -1);
    Buckets B;
    bool built = false;
    std::vector<std::pair<double, int>> cand;
    for (int c = 0; c < N; ++c) {
      int nv = m.nvrt[c], base = m.lsq_ptr[c];
      int gap[4], ngap = 0;
      for (int in = 0; in < nv; ++in) {
        int nb = m.nghbre[4 * c + (in + nv - 2) % nv];
        if (nb == -1) gap[ngap++] = in;
        else m.lsq_cell[base + in] = nb;
      }
      if (ngap > 2) return fmt("LSQ face-neighbour stencil: cell %ld has more than 2 boundary edges", c);
      if (ngap == 0) continue;
      if (!built) { B.build(m.cx, m.cy); built = true; }
      knearest(B, m.cx, m.cy, m.cx[c], m.cy[c], 8, cand);
      int filled = 0;
      for (auto &pr : cand) {
        int j = pr.second;
        if (j == c) continue;
        bool isnb = false;
        for (int e = 0; e < nv; ++e) isnb = isnb || m.nghbre[4 * c + e] == j;
        if (!isnb) m.lsq_cell[base + gap[filled++]] = j;
        if (filled == ngap) break;
      }
      if (filled < ngap) return fmt("LSQ face-neighbour stencil: too few candidates for cell %ld", c);
    }
  } else {
    // setup_nn (src/gradient_lsq.f90:227-271): cells sharing a vertex,
    // ascending and unique
    std::vector<int> tmp;
    m.lsq_cell.clear();
    m.lsq_cell.reserve(size_t(N) * 13);
    for (int c = 0; c < N; ++c) {
      tmp.clear();
      for (int k = 0; k < m.nvrt[c]; ++k) {
        int v = m.cnodec[4 * c + k];
        for (int q = m.n2c_ptr[v]; q < m.n2c_ptr[v + 1]; ++q)
          if (m.n2c[q] != c) tmp.push_back(m.n2c[q]);
      }
      std::sort(tmp.begin(), tmp.end());
      tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
      m.lsq_cell.insert(m.lsq_cell.end(), tmp.begin(), tmp.end());
      m.lsq_ptr[c + 1] = m.lsq_ptr[c] + int(tmp.size());
    }
  }
  // weights w = 1/d^p and coef = G^-1 D^T, D rows w*(dx,dy):
  // src/gradient_lsq.f90:137-201 and 277-348 (same arithmetic in both)
  const int nnz = m.lsq_ptr[N];
  m.lsq_w.assign(nnz, 0.0);
  m.lsq_coef.assign(2 * size_t(nnz), 0.0);
  std::vector<double> d1, d2;
  for (int c = 0; c < N; ++c) {
    int b = m.lsq_ptr[c], n = m.lsq_ptr[c + 1] - b;
    d1.assign(n, 0.0);
    d2.assign(n, 0.0);
    for (int k = 0; k < n; ++k) {
      int j = m.lsq_cell[b + k];
      double sx, sy;
      pair_shift(m, c, j, sx, sy);
      double xj = m.cx[j] + sx, yj = m.cy[j] + sy;
      double dis = std::sqrt((xj - m.cx[c]) * (xj - m.cx[c]) + (yj - m.cy[c]) * (yj - m.cy[c]));
      double w = 0.0;
      if (dis > 0.0) {
        w = 1.0 / std::pow(dis, p);
        m.lsq_w[b + k] = w;
      }
      d1[k] = w * (xj - m.cx[c]);
      d2[k] = w * (yj - m.cy[c]);
    }
    double g11 = 0, g12 = 0, g21 = 0, g22 = 0;
    for (int k = 0; k < n; ++k) g11 = g11 + d1[k] * d1[k];
    for (int k = 0; k < n; ++k) g12 = g12 + d1[k] * d2[k];
    for (int k = 0; k < n; ++k) g21 = g21 + d2[k] * d1[k];
    for (int k = 0; k < n; ++k) g22 = g22 + d2[k] * d2[k];
    double det = g11 * g22 - g12 * g21;
    double i11 = 1.0 / det * g22, i22 = 1.0 / det * g11, i12 = -(1.0 / det * g12), i21 = -(1.0 / det * g21);
    for (int k = 0; k < n; ++k) {
      double a = 0.0, bb = 0.0;
      a = a + i11 * d1[k];
      a = a + i12 * d2[k];
      bb = bb + i21 * d1[k];
      bb = bb + i22 * d2[k];
      m.lsq_coef[2 * size_t(b + k)] = a;
      m.lsq_coef[2 * size_t(b + k) + 1] = bb;
    }
  }
  // exactness for f = 2x + y to 1e-10 (grad_lsq_verify, src/gradient_lsq.f90:490-529)
  for (int c = 0; c < N; ++c) {
    double fx = 0.0, fy = 0.0;
    for (int k = m.lsq_ptr[c]; k < m.lsq_ptr[c + 1]; ++k) {
      int j = m.lsq_cell[k];
      double sx, sy;
      pair_shift(m, c, j, sx, sy);
      double diff = 1.0 * (m.cy[j] + sy) + 2.0 * (m.cx[j] + sx) - (1.0 * m.cy[c] + 2.0 * m.cx[c]);
      fx = fx + m.lsq_coef[2 * size_t(k)] * diff * m.lsq_w[k];
      fy = fy + m.lsq_coef[2 * size_t(k) + 1] * diff * m.lsq_w[k];
    }
    if (std::fabs(fx - 2.0) > 1e-10 || std::fabs(fy - 1.0) > 1e-10)
      return fmt("LSQ coefficients are not exact for a linear function at cell %ld", c);
  }
  return "";
}

void ggcb_setup(MeshPrep &m) {
  // src/gradient_ggcb.f90:48-110: face value weighted by the distances of the
  // face centre to both cell centres
  const int N = m.ncells;
  m.ggcb_ptr.assign(4 * N, -1);
  m.ggcb_coef0.assign(2 * N, 0.0);
  m.ggcb_coefnb.assign(8 * N, 0.0);
  for (int c = 0; c < N; ++c)
    for (int e = 0; e < m.nvrt[c]; ++e) {
      int je = m.cedge[4 * c + e];
      int c1 = m.ec1[je], c2 = m.ec2[je];
      double af = m.ea[je], nxf = m.enx[je] * m.nrmlsign[4 * c + e], nyf = m.eny[je] * m.nrmlsign[4 * c + e];
      double xf = m.ex[je], yf = m.ey[je], xc, yc;
      side_centroid(m, je, c, xc, yc);
      double dx = xf - xc, dy = yf - yc;
      double d0 = std::sqrt(dx * dx + dy * dy);
      double xo = xc, yo = yc;
      int other = c;
      if (c1 >= 0 && c1 != c) other = c1;
      else if (c2 >= 0 && c2 != c) other = c2;
      if (other != c) side_centroid(m, je, other, xo, yo);
      m.ggcb_ptr[4 * c + e] = other;
      dx = xf - xo;
      dy = yf - yo;
      double d1 = std::sqrt(dx * dx + dy * dy);
      m.ggcb_coef0[2 * c] = m.ggcb_coef0[2 * c] + d1 / (d0 + d1) * nxf * af;
      m.ggcb_coef0[2 * c + 1] = m.ggcb_coef0[2 * c + 1] + d1 / (d0 + d1) * nyf * af;
      m.ggcb_coefnb[8 * c + 2 * e] = d0 / (d0 + d1) * nxf * af;
      m.ggcb_coefnb[8 * c + 2 * e + 1] = d0 / (d0 + d1) * nyf * af;
    }
}

void ggnb_setup(MeshPrep &m) {
  // src/gradient_ggnb.f90:217-345: Green-Gauss over node values that are
  // inverse-distance averages of the cells around each node.  The reference
  // folds the node values into cell-to-cell coefficients (its coef0 is the
  // cell's own share of its node values); kept here as the two factors:
  // per node incidence idw(v)/d(cell, v), per cell and node coefedg.
  const int N = m.ncells, V = m.nnodes;
  std::vector<double> idw(V, 0.0);
  for (int v = 0; v < V; ++v) {
    if (m.n2c_ptr[v + 1] == m.n2c_ptr[v]) continue;
    double idt = 0.0;
    for (int q = m.n2c_ptr[v]; q < m.n2c_ptr[v + 1]; ++q) {
      int c = m.n2c[q];
      double xv, yv;
      node_in_cell(m, c, v, xv, yv);
      double dx = m.cx[c] - xv, dy = m.cy[c] - yv;
      idt = idt + 1.0 / std::sqrt(dx * dx + dy * dy);
    }
    idw[v] = 1.0 / idt;
  }
  m.ggnb_ncoef.assign(m.n2c.size(), 0.0);
  for (int v = 0; v < V; ++v)
    for (int q = m.n2c_ptr[v]; q < m.n2c_ptr[v + 1]; ++q) {
      int c = m.n2c[q];
      double xv, yv;
      node_in_cell(m, c, v, xv, yv);
      double dx = m.cx[c] - xv, dy = m.cy[c] - yv;
      m.ggnb_ncoef[q] = idw[v] / std::sqrt(dx * dx + dy * dy);
    }
  // coefedg(in) = sum over the cell's edges touching node in of af*n/2,
  // edge ends visited n1 then n2 (src/gradient_ggnb.f90:321-342)
  m.ggnb_edg.assign(8 * size_t(N), 0.0);
  for (int c = 0; c < N; ++c)
    for (int in = 0; in < m.nvrt[c]; ++in) {
      int iv = m.cnodec[4 * c + in];
      for (int e = 0; e < m.nvrt[c]; ++e) {
        int je = m.cedge[4 * c + e];
        for (int j = 0; j < 2; ++j) {
          int jv = m.canon[j == 0 ? m.en1[je] : m.en2[je]];
          if (jv != iv) continue;
          double af = m.ea[je], nxf = m.enx[je] * m.nrmlsign[4 * c + e], nyf = m.eny[je] * m.nrmlsign[4 * c + e];
          m.ggnb_edg[8 * size_t(c) + 2 * in] = m.ggnb_edg[8 * size_t(c) + 2 * in] + af * nxf / 2.0;
          m.ggnb_edg[8 * size_t(c) + 2 * in + 1] = m.ggnb_edg[8 * size_t(c) + 2 * in + 1] + af * nyf / 2.0;
        }
      }
    }
}

}  // namespace

std::string build_mesh(MeshPrep &m, int nnodes, const double *xy, int ntri, int nquad, const int *cell_nodes,
                       const int *node_canon, int nb, const int *bnd_ncells, const int *bnd_types,
                       const int *bnd_cells, const SchemeFlags &flags) {
  if (nnodes <= 0 || ntri < 0 || nquad < 0 || ntri + nquad <= 0) return "empty mesh";
  if (nb < 0) return "negative boundary count";
  m = MeshPrep();
  m.nnodes = nnodes;
  m.ntri = ntri;
  m.nquad = nquad;
  m.ncells = ntri + nquad;
  m.periodic = node_canon != nullptr;
  m.nxy.assign(xy, xy + 2 * size_t(nnodes));
  m.canon.resize(nnodes);
  for (int v = 0; v < nnodes; ++v) {
    m.canon[v] = node_canon ? node_canon[v] : v;
    if (m.canon[v] < 0 || m.canon[v] >= nnodes) return fmt("node %ld has canonical id %ld", v, m.canon[v]);
  }
  const int N = m.ncells;
  m.nvrt.assign(N, 0);
  m.cnode.assign(4 * size_t(N), -1);
  m.cnodec.assign(4 * size_t(N), -1);
  const int *p = cell_nodes;
  for (int c = 0; c < N; ++c) {
    m.nvrt[c] = c < ntri ? 3 : 4;
    for (int k = 0; k < m.nvrt[c]; ++k) {
      int v = *p++;
      if (v < 0 || v >= nnodes) return fmt("cell %ld references node %ld outside the mesh", c, v);
      m.cnode[4 * size_t(c) + k] = v;
      m.cnodec[4 * size_t(c) + k] = m.canon[v];
    }
  }
  m.nb = nb;
  m.bnd_type.assign(bnd_types, bnd_types + nb);
  m.bnd_ncells.assign(bnd_ncells, bnd_ncells + nb);
  m.bnd_cptr.assign(nb + 1, 0);
  for (int b = 0; b < nb; ++b) {
    if (bnd_ncells[b] < 0) return "negative boundary cell count";
    if (bnd_types[b] < 0 || bnd_types[b] > 2)
      return fmt("boundary %ld: type %ld not implemented (freestream, slip_wall, dirichlet)", b + 1, bnd_types[b]);
    m.bnd_cptr[b + 1] = m.bnd_cptr[b] + bnd_ncells[b];
  }
  m.bnd_cells.assign(bnd_cells, bnd_cells + m.bnd_cptr[nb]);
  std::string err = topology(m);
  if (!err.empty()) return err;
  if (flags.grad_method < 1 || flags.grad_method > 3) return "grad_method must be 1 (GGCB), 2 (GGNB) or 3 (LSQ)";
  if (flags.grad_method == 3 || flags.limiter > 0) {
    err = lsq_stencil(m, flags.grad_method == 3 ? flags.lsq_nn != 0 : true, flags.grad_method == 3 ? flags.lsq_p : 0.0);
    if (!err.empty()) return err;
  }
  if (flags.grad_method == 1) ggcb_setup(m);
  if (flags.grad_method == 2) ggnb_setup(m);
  return "";
}

std::string build_mesh_batch(MeshPrep &m, int ninst, const int *cell_offsets, int nnodes, const double *xy, int ntri,
                             int nquad, const int *cell_nodes, int nb, const int *bnd_ncells, const int *bnd_types,
                             const int *bnd_cells, const SchemeFlags &flags, int nthreads) {
  const int N = ntri + nquad;
  if (ninst <= 0 || N <= 0 || nnodes <= 0) return "empty batch";
  if (ntri > 0 && nquad > 0) return "per-instance preparation needs cells of one kind";
  if (nb < 0) return "negative boundary count";
  const int nv = ntri > 0 ? 3 : 4;
  // node range of every instance; instances own consecutive ranges
  std::vector<int> nlo(ninst), nhi(ninst);
  for (int k = 0; k < ninst; ++k) {
    int lo = nnodes, hi = -1;
    for (long q = long(cell_offsets[k]) * nv; q < long(cell_offsets[k + 1]) * nv; ++q) {
      const int v = cell_nodes[q];
      if (v < 0 || v >= nnodes) return fmt("cell %ld references node %ld outside the mesh", q / nv, v);
      lo = std::min(lo, v);
      hi = std::max(hi, v);
    }
    nlo[k] = lo;
    nhi[k] = hi;
    if (lo != (k == 0 ? 0 : nhi[k - 1] + 1)) return "instances do not own consecutive node ranges";
  }
  if (nhi[ninst - 1] != nnodes - 1) return "instances do not own consecutive node ranges";
  // every boundary's cells, split by instance: a boundary may span several
  // instances (same-named boundaries concatenated), its cells listed in
  // instance order; instance k gets the boundaries it has cells on, in order
  std::vector<int> bptr(nb + 1, 0);
  for (int b = 0; b < nb; ++b) {
    if (bnd_ncells[b] < 0) return "negative boundary cell count";
    if (bnd_types[b] < 0 || bnd_types[b] > 2)
      return fmt("boundary %ld: type %ld not implemented (freestream, slip_wall, dirichlet)", b + 1, bnd_types[b]);
    bptr[b + 1] = bptr[b] + bnd_ncells[b];
  }
  // sub[b]: for each instance with cells on boundary b, (instance, first,
  // end) positions in the boundary's list
  std::vector<std::vector<std::array<int, 3>>> sub(nb);
  std::vector<std::vector<int>> inst_b(ninst);  // boundaries of instance k
  for (int b = 0; b < nb; ++b) {
    int prev = -1;
    for (int i = bptr[b]; i < bptr[b + 1]; ++i) {
      const int c = bnd_cells[i];
      if (c < 0 || c >= N) return fmt("boundary %ld lists cell %ld outside the mesh", b + 1, c);
      const int k = int(std::upper_bound(cell_offsets, cell_offsets + ninst + 1, c) - cell_offsets) - 1;
      if (k < prev) return fmt("boundary %ld does not list its cells instance by instance", b + 1, 0);
      if (k != prev) {
        sub[b].push_back({k, i, i});
        inst_b[k].push_back(b);
        prev = k;
      }
      sub[b].back()[2] = i + 1;
    }
  }
  // every instance on its own, on nthreads workers
  std::vector<MeshPrep> part(ninst);
  std::vector<std::string> errs(ninst);
  std::atomic<int> next(0);
  auto work = [&]() {
    for (int k = next++; k < ninst; k = next++) {
      const int c0 = cell_offsets[k], nc = cell_offsets[k + 1] - c0, v0 = nlo[k], nn = nhi[k] - v0 + 1;
      std::vector<int> cn(size_t(nc) * nv);
      for (size_t q = 0; q < cn.size(); ++q) cn[q] = cell_nodes[size_t(c0) * nv + q] - v0;
      const int nbk = int(inst_b[k].size());
      std::vector<int> bn(nbk), bt(nbk), bc;
      for (int q = 0; q < nbk; ++q) {
        const int b = inst_b[k][q];
        for (const auto &r : sub[b])
          if (r[0] == k) {
            bn[q] = r[2] - r[1];
            for (int i = r[1]; i < r[2]; ++i) bc.push_back(bnd_cells[i] - c0);
          }
        bt[q] = bnd_types[b];
      }
      errs[k] = build_mesh(part[k], nn, xy + 2 * size_t(v0), ntri > 0 ? nc : 0, ntri > 0 ? 0 : nc, cn.data(),
                           nullptr, nbk, bn.data(), bt.data(), bc.data(), flags);
    }
  };
  run_threads(std::max(1, std::min(nthreads, ninst)), [&](int) { work(); });
  // every failing instance is named (callers may drop them all at once)
  {
    std::string bad, first;
    for (int k = 0; k < ninst; ++k)
      if (!errs[k].empty()) {
        bad += (bad.empty() ? "" : ",") + std::to_string(k);
        if (first.empty()) first = errs[k];
      }
    if (!bad.empty()) return "instances [" + bad + "] failed; the first: " + first;
  }
  // merge in instance order: cells, nodes and edges of instance k follow
  // those of instance k - 1, the numbering a whole-concatenation build gives
  m = MeshPrep();
  m.nnodes = nnodes;
  m.ntri = ntri;
  m.nquad = nquad;
  m.ncells = N;
  m.nb = nb;
  std::vector<int> cofs(ninst + 1, 0), vofs(ninst + 1, 0), eofs(ninst + 1, 0);
  for (int k = 0; k < ninst; ++k) {
    cofs[k + 1] = cofs[k] + part[k].ncells;
    vofs[k + 1] = vofs[k] + part[k].nnodes;
    eofs[k + 1] = eofs[k] + part[k].nedges;
  }
  // boundaries: each one's cells and edges, instance by instance (the order
  // a whole-concatenation build walks its listed cells in)
  m.bnd_type.assign(bnd_types, bnd_types + nb);
  m.bnd_ncells.assign(nb, 0);
  m.bnd_nedges.assign(nb, 0);
  m.bnd_cptr.assign(1, 0);
  m.bnd_eptr.assign(1, 0);
  for (int b = 0; b < nb; ++b) {
    for (const auto &r : sub[b]) {
      const int k = r[0];
      const MeshPrep &p = part[k];
      const int q = int(std::find(inst_b[k].begin(), inst_b[k].end(), b) - inst_b[k].begin());
      for (int i = p.bnd_cptr[q]; i < p.bnd_cptr[q + 1]; ++i) m.bnd_cells.push_back(p.bnd_cells[i] + cofs[k]);
      for (int i = p.bnd_eptr[q]; i < p.bnd_eptr[q + 1]; ++i) m.bnd_edges.push_back(p.bnd_edges[i] + eofs[k]);
      m.bnd_ncells[b] += p.bnd_ncells[q];
      m.bnd_nedges[b] += p.bnd_nedges[q];
    }
    m.bnd_cptr.push_back(int(m.bnd_cells.size()));
    m.bnd_eptr.push_back(int(m.bnd_edges.size()));
  }
  // field by field: every part's slice at its offset (node, cell and edge ids
  // shifted, pointer arrays continued), the parts copied on the worker
  // threads, each field of the parts released once it is merged
  auto run_parts = [&](auto &&fn) {
    std::atomic<int> next{0};
    run_threads(std::max(1, std::min(nthreads, ninst)), [&](int) {
      for (int k = next++; k < ninst; k = next++) fn(k);
    });
  };
  auto starts = [&](auto field) {
    std::vector<size_t> st(ninst + 1, 0);
    for (int k = 0; k < ninst; ++k) st[k + 1] = st[k] + (part[k].*field).size();
    return st;
  };
  auto merge_copy = [&](auto field) {
    auto &dst = m.*field;
    const std::vector<size_t> st = starts(field);
    dst.resize(st[ninst]);
    run_parts([&](int k) {
      auto &src = part[k].*field;
      std::copy(src.begin(), src.end(), dst.begin() + st[k]);
      std::remove_reference_t<decltype(src)>().swap(src);
    });
  };
  auto merge_ids = [&](auto field, const std::vector<int> &add) {
    auto &dst = m.*field;
    const std::vector<size_t> st = starts(field);
    dst.resize(st[ninst]);
    run_parts([&](int k) {
      auto &src = part[k].*field;
      for (size_t q = 0; q < src.size(); ++q) dst[st[k] + q] = src[q] >= 0 ? src[q] + add[k] : src[q];
      std::vector<int>().swap(src);
    });
  };
  // CSR pointer arrays (each part's starting at 0): one leading 0, then each
  // part's entries after its first plus the running total before it
  auto merge_ptr = [&](auto field) {
    auto &dst = m.*field;
    std::vector<size_t> st(ninst + 1, 1);
    std::vector<int> base(ninst + 1, 0);
    bool any = false;
    for (int k = 0; k < ninst; ++k) {
      const auto &src = part[k].*field;
      const size_t n = src.empty() ? 0 : src.size() - 1;
      any = any || !src.empty();
      st[k + 1] = st[k] + n;
      base[k + 1] = base[k] + (src.empty() ? 0 : src.back());
    }
    if (!any) return;
    dst.assign(st[ninst], 0);
    run_parts([&](int k) {
      auto &src = part[k].*field;
      for (size_t q = 1; q < src.size(); ++q) dst[st[k] + q - 1] = src[q] + base[k];
      std::vector<int>().swap(src);
    });
  };
  merge_copy(&MeshPrep::nxy);
  merge_ids(&MeshPrep::canon, vofs);
  merge_copy(&MeshPrep::nvrt);
  merge_ids(&MeshPrep::cnode, vofs);
  merge_ids(&MeshPrep::cnodec, vofs);
  merge_ids(&MeshPrep::nghbre, cofs);
  merge_ids(&MeshPrep::cedge, eofs);
  merge_copy(&MeshPrep::cx);
  merge_copy(&MeshPrep::cy);
  merge_copy(&MeshPrep::vol);
  merge_copy(&MeshPrep::nrmlsign);
  merge_ptr(&MeshPrep::n2c_ptr);
  merge_ids(&MeshPrep::n2c, cofs);
  merge_ids(&MeshPrep::en1, vofs);
  merge_ids(&MeshPrep::en2, vofs);
  merge_ids(&MeshPrep::ec1, cofs);
  merge_ids(&MeshPrep::ec2, cofs);
  merge_copy(&MeshPrep::ex);
  merge_copy(&MeshPrep::ey);
  merge_copy(&MeshPrep::ea);
  merge_copy(&MeshPrep::enx);
  merge_copy(&MeshPrep::eny);
  merge_copy(&MeshPrep::esx);
  merge_copy(&MeshPrep::esy);
  merge_ids(&MeshPrep::edge_intr, eofs);
  merge_ids(&MeshPrep::cell_intr, cofs);
  for (int k = 0; k < ninst; ++k) {
    m.ncells_bndr += part[k].ncells_bndr;
    m.nedges_bndr += part[k].nedges_bndr;
  }
  merge_ptr(&MeshPrep::lsq_ptr);
  merge_ids(&MeshPrep::lsq_cell, cofs);
  merge_copy(&MeshPrep::lsq_w);
  merge_copy(&MeshPrep::lsq_coef);
  merge_ids(&MeshPrep::ggcb_ptr, cofs);
  merge_copy(&MeshPrep::ggcb_coef0);
  merge_copy(&MeshPrep::ggcb_coefnb);
  merge_copy(&MeshPrep::ggnb_ncoef);
  merge_copy(&MeshPrep::ggnb_edg);
  m.nedges = eofs[ninst];
  return "";
}

void run_threads(int nt, const std::function<void(int)> &body) {
  std::vector<std::exception_ptr> err(std::max(nt, 1));
  auto guarded = [&](int t) {
    try {
      body(t);
    } catch (...) {
      err[t] = std::current_exception();
    }
  };
  std::vector<std::thread> th;
  int started = 1;
  try {
    th.reserve(std::max(nt - 1, 0));
    for (; started < nt; ++started) th.emplace_back(guarded, started);
  } catch (...) {
    // no thread for parts started .. nt - 1: they run below
  }
  for (int t = started; t < nt; ++t) guarded(t);
  guarded(0);
  for (auto &x : th) x.join();
  for (auto &e : err)
    if (e) std::rethrow_exception(e);
}

}  // namespace fvs
