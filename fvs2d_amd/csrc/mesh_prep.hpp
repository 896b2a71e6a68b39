// Host-side mesh preparation: topology, geometry and gradient stencils with
// the numbering and orientation conventions of src/grid_procs.f90:170-794 and
// the stencil constructions of src/gradient_lsq.f90, src/gradient_ggcb.f90 and
// src/gradient_ggnb.f90.  All ids are 0-based, -1 meaning "none".
#pragma once
#include <functional>
#include <string>
#include <vector>

namespace fvs {

struct SchemeFlags {
  int grad_method = 3;  // 1 GGCB, 2 GGNB, 3 LSQ
  int lsq_nn = 1;
  double lsq_p = 0.0;
  int limiter = 0;
};

struct MeshPrep {
  int nnodes = 0, ncells = 0, ntri = 0, nquad = 0, nedges = 0;
  bool periodic = false;
  std::vector<double> nxy;   // 2*nnodes
  std::vector<int> canon;    // per node (identity when not periodic)
  std::vector<int> nvrt;     // ncells
  std::vector<int> cnode;    // 4*ncells physical ids (-1 pad)
  std::vector<int> cnodec;   // 4*ncells canonical ids
  std::vector<int> nghbre;   // 4*ncells: neighbour across local edge e=(v_e,v_e+1)
  std::vector<int> cedge;    // 4*ncells: global edge id of local edge e
  std::vector<double> cx, cy, vol, nrmlsign;
  std::vector<int> n2c_ptr, n2c;  // canonical node -> cells (ascending)
  // edges
  std::vector<int> en1, en2, ec1, ec2;
  std::vector<double> ex, ey, ea, enx, eny, esx, esy;
  std::vector<int> edge_intr, cell_intr;
  int ncells_bndr = 0, nedges_bndr = 0;
  // boundaries (the .bc file + derived edge lists)
  int nb = 0;
  std::vector<int> bnd_type, bnd_ncells, bnd_cptr, bnd_cells, bnd_nedges, bnd_eptr, bnd_edges;
  // LSQ stencil (also the limiter's stencil); for GG gradients with a limiter
  // this is the node-neighbour stencil
  std::vector<int> lsq_ptr, lsq_cell;
  std::vector<double> lsq_w, lsq_coef;  // coef: 2 per entry
  // GGCB: per local edge (ptr, coefnb(2)); coef0(2) per cell
  std::vector<int> ggcb_ptr;
  std::vector<double> ggcb_coef0, ggcb_coefnb;
  // GGNB: per node incidence (aligned with n2c) idw(v)/d(cell, v); per cell
  // and local node (4 slots) coefedg (2)
  std::vector<double> ggnb_ncoef, ggnb_edg;
};

// Builds everything; returns "" on success or an error message.
std::string build_mesh(MeshPrep &m, int nnodes, const double *xy, int ntri, int nquad, const int *cell_nodes,
                       const int *node_canon, int nb, const int *bnd_ncells, const int *bnd_types,
                       const int *bnd_cells, const SchemeFlags &flags);

// The same for a batch of ninst independent instances concatenated in
// instance order (instance k owns cells [cell_offsets[k], cell_offsets[k+1]),
// a consecutive node range and consecutive boundaries, cells of one kind):
// each instance is prepared on its own, on up to nthreads threads, and the
// results are merged with node, cell and edge offsets, which is the numbering
// the whole-concatenation build gives (stencil searches stay inside an
// instance).  Returns "" or an error message.
std::string build_mesh_batch(MeshPrep &m, int ninst, const int *cell_offsets, int nnodes, const double *xy, int ntri,
                             int nquad, const int *cell_nodes, int nb, const int *bnd_ncells, const int *bnd_types,
                             const int *bnd_cells, const SchemeFlags &flags, int nthreads);

// Runs body(0) .. body(nt - 1): body(0) on the calling thread, the others on
// worker threads; a part whose thread cannot be started runs on the calling
// thread instead.  Every thread is joined before the first exception a part
// threw (an allocation that failed) is rethrown to the caller, so no
// exception ends a worker thread or leaves one running.
void run_threads(int nt, const std::function<void(int)> &body);

// Frame helpers for periodic meshes (no reference counterpart): the centroid
// of edge `je`'s cell `c` expressed in the edge's c1 frame.
inline void side_centroid(const MeshPrep &m, int je, int c, double &x, double &y) {
  if (c == m.ec1[je]) { x = m.cx[c]; y = m.cy[c]; }
  else { x = m.cx[c] + m.esx[je]; y = m.cy[c] + m.esy[je]; }
}

}  // namespace fvs
