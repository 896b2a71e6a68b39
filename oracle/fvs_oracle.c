/*
 * fvs_oracle.c -- CPU restatement of the fvs2d residual / time-step path.
 *
 * TEST INFRASTRUCTURE ONLY (the results the tests compare with, and the timed
 * CPU baseline).  The
 * product never links, loads or calls this file.
 *
 * It restates, loop for loop and in the reference's summation order, the
 * Fortran routines cited beside each function (shirzadgit/fvs2d, src/...).
 * It is deliberately plain C: no reassociation, no FMA contraction (build
 * with -ffp-contract=off), 0-based indices with -1 for "no cell".  Serial by
 * default; orc_set_threads runs the per-cell loops on OpenMP threads with the
 * same per-cell arithmetic and order (bitwise the serial results).
 *
 * Builder-new extensions (no reference counterpart, marked as such):
 *   - periodic node identification (node_canon): topology on canonical node
 *     ids, geometry in each cell's own frame, a per-edge shift carrying the
 *     second cell's centroid into the first cell's frame;
 *   - SSP-RK3 (Shu-Osher form).
 *
 * Pinning status of each function (what checks it) is recorded in the header
 * comment of the test that checks it; functions are never checked by
 * re-deriving the same formula.
 */
#include "fvs_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
/* per-cell loops run on nthr threads when orc_set_threads asked for more
 * than one; every cell's arithmetic is the serial one */
#define ORC_PAR _Pragma("omp parallel for if(o->nthr > 1) num_threads(o->nthr > 1 ? o->nthr : 1) schedule(static)")

/* Planted mistakes, one per build, for the test that shows each pin fails on
 * a plausible error (tests/test_oracle_mutants.py compiles throw-away
 * variants with -DORC_MUTANT=k into a temporary directory; the oracle library
 * itself is built with ORC_MUTANT 0, i.e. none of them). */
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif
#define MUT_GGCB_SWAP_WEIGHTS 1   /* d0/(d0+d1) on the cell, d1/(d0+d1) on the neighbour */
#define MUT_GGNB_IDW_SQUARE 2     /* node weights 1/d^2 instead of 1/d */
#define MUT_GGNB_SELF_TWICE 3     /* the self incidence kept in coefnb as well as in coef0 */
#define MUT_LIM_FACE_STENCIL 4    /* limiter bounds over face neighbours only */
#define MUT_LIM_GRAD_SWAP 5       /* face extrapolation with d/dx and d/dy exchanged */
#define MUT_LIM_RHO_ONLY 6        /* phi of the density alone instead of the min over variables */
#define MUT_SLIP_NO_FACTOR2 7     /* slip-wall ghost with zero instead of reversed normal velocity */
#define MUT_RK2_SWAP_BC 8         /* order-2 table with b and c exchanged */
#define MUT_RK1_SWAP_AD 9         /* order-1 table with a and d exchanged */
#define MUT_RK3_H3 10             /* third classical stage advanced by dt/2 instead of dt */

static char g_err[512];
const char *orc_last_error(void) { return g_err; }
static void set_err(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

#define NV 4
static const double PI = 3.14159265358979323846; /* acos(-1.d0) */

struct orc {
  orc_opts op;
  int nnodes, ncells, ntri, nquad, nedges;
  const double *dummy;
  double *nxy;  /* physical node coordinates */
  int *canon;   /* canonical node id per physical node */
  int *nvrt;
  int *cnode;   /* 4 per cell, physical node ids */
  int *cnodec;  /* 4 per cell, canonical node ids */
  int *nghbr, *nghbre, *cedge;
  double *cx, *cy, *vol, *nrmlsign;
  int *n2c_ptr, *n2c; /* canonical node -> cells */
  int *en1, *en2, *ec1, *ec2;
  double *ex, *ey, *ea, *enx, *eny, *esx, *esy;
  int nedges_intr, nedges_bndr, ncells_intr, ncells_bndr;
  int *edge_intr, *cell_intr;
  int nb;
  int *bnd_type, *bnd_ncells, *bnd_cptr, *bnd_cells, *bnd_nedges, *bnd_eptr, *bnd_edges;
  /* LSQ stencil (CSR) */
  int *lsq_ptr, *lsq_cell;
  double *lsq_w, *lsq_coef;
  /* GGCB */
  int *ggcb_ptr;
  double *ggcb_coef0, *ggcb_coefnb;
  /* GGNB */
  double *ggnb_coef0, *ggnb_coefedg;
  int *ggnb_ptr, *ggnb_cell, *ggnb_in;
  double *ggnb_coefnb;
  /* solution */
  double *pvar, *cvar, *grad, *resid, *phi, *ws, *dt_local;
  double pvar_inf[4];
  double kappa;
  /* threaded evaluation (orc_set_threads): per cell, its faces in the serial
   * scatter order as signed (face + 1) -- interior faces by position in
   * edge_intr, then boundary faces after them -- and the face products */
  int nthr;
  int *cf_ptr, *cf;
  double *fa, *wa;
};

static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s);
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
  return p;
}

/* cosd/sind of the Fortran reference (degrees) */
static double cosd(double x) { return cos(x * (PI / 180.0)); }
static double sind(double x) { return sin(x * (PI / 180.0)); }

/* coordinates of canonical node v as seen from cell c (its own frame) */
static void cell_node_xy(const orc_t *o, int c, int v, double *x, double *y) {
  for (int k = 0; k < o->nvrt[c]; ++k)
    if (o->cnodec[4 * c + k] == v) {
      int p = o->cnode[4 * c + k];
      *x = o->nxy[2 * p];
      *y = o->nxy[2 * p + 1];
      return;
    }
  *x = NAN; *y = NAN;
}

/* shift carrying cell j's centroid into cell i's frame through a shared
 * canonical node (zero for non-periodic meshes) -- builder-new */
static void pair_shift(const orc_t *o, int i, int j, double *sx, double *sy) {
  *sx = 0.0; *sy = 0.0;
  if (!o->canon) return;
  for (int a = 0; a < o->nvrt[i]; ++a)
    for (int b = 0; b < o->nvrt[j]; ++b)
      if (o->cnodec[4 * i + a] == o->cnodec[4 * j + b]) {
        int pi_ = o->cnode[4 * i + a], pj = o->cnode[4 * j + b];
        *sx = o->nxy[2 * pi_] - o->nxy[2 * pj];
        *sy = o->nxy[2 * pi_ + 1] - o->nxy[2 * pj + 1];
        return;
      }
}

/* src/grid_procs.f90:800-808 */
static double tri_area(double x1, double x2, double x3, double y1, double y2, double y3) {
  return 0.5 * (x1 * (y2 - y3) + x2 * (y3 - y1) + x3 * (y1 - y2));
}

/* ------------------------------------------------------------------------ */
/* grid_data: src/grid_procs.f90:170-794                                    */
/* ------------------------------------------------------------------------ */
static int grid_data(orc_t *o) {
  int N = o->ncells;
  /* centroids 185-208 */
  for (int ic = 0; ic < N; ++ic) {
    double xc = 0.0, yc = 0.0;
    for (int iv = 0; iv < o->nvrt[ic]; ++iv) {
      int vp = o->cnode[4 * ic + iv];
      xc = xc + o->nxy[2 * vp];
      yc = yc + o->nxy[2 * vp + 1];
    }
    double d = (o->nvrt[ic] == 3) ? 3.0 : 4.0;
    o->cx[ic] = xc / d;
    o->cy[ic] = yc / d;
  }
  /* volumes 214-227 */
  for (int ic = 0; ic < N; ++ic) {
    const int *nd = &o->cnode[4 * ic];
    double x1 = o->nxy[2 * nd[0]], y1 = o->nxy[2 * nd[0] + 1];
    double x2 = o->nxy[2 * nd[1]], y2 = o->nxy[2 * nd[1] + 1];
    double x3 = o->nxy[2 * nd[2]], y3 = o->nxy[2 * nd[2] + 1];
    if (o->nvrt[ic] == 3) {
      o->vol[ic] = tri_area(x1, x2, x3, y1, y2, y3);
    } else {
      double x4 = o->nxy[2 * nd[3]], y4 = o->nxy[2 * nd[3] + 1];
      o->vol[ic] = tri_area(x1, x2, x3, y1, y2, y3) + tri_area(x1, x3, x4, y1, y3, y4);
    }
  }
  /* node -> cell, 273-296 (canonical nodes) */
  int NN = o->nnodes;
  o->n2c_ptr = xcalloc(NN + 1, sizeof(int));
  for (int ic = 0; ic < N; ++ic)
    for (int iv = 0; iv < o->nvrt[ic]; ++iv) o->n2c_ptr[o->cnodec[4 * ic + iv] + 1]++;
  for (int i = 0; i < NN; ++i) o->n2c_ptr[i + 1] += o->n2c_ptr[i];
  o->n2c = xcalloc(o->n2c_ptr[NN], sizeof(int));
  int *fill = xcalloc(NN, sizeof(int));
  for (int ic = 0; ic < N; ++ic)
    for (int iv = 0; iv < o->nvrt[ic]; ++iv) {
      int v = o->cnodec[4 * ic + iv];
      o->n2c[o->n2c_ptr[v] + fill[v]++] = ic;
    }
  free(fill);
  /* neighbours, 325-366 */
  for (int ic = 0; ic < N; ++ic) {
    int nv = o->nvrt[ic];
    for (int iv = 0; iv < nv; ++iv) {
      int vL = (iv < nv - 1) ? o->cnodec[4 * ic + iv + 1] : o->cnodec[4 * ic];
      int vR = o->cnodec[4 * ic + iv];
      int found = 0, jc = -1, im = 0;
      for (int j = o->n2c_ptr[vR]; j < o->n2c_ptr[vR + 1] && !found; ++j) {
        jc = o->n2c[j];
        int nj = o->nvrt[jc];
        for (int ii = 0; ii < nj; ++ii) {
          int v2 = (ii > 0) ? o->cnodec[4 * jc + ii - 1] : o->cnodec[4 * jc + nj - 1];
          int v1 = o->cnodec[4 * jc + ii];
          if (v1 == vR && v2 == vL) {
            found = 1;
            im = ii + 1;
            if (im > nj - 1) im -= nj;
            break;
          }
        }
      }
      int in = iv + 2;
      if (in > nv - 1) in -= nv;
      if (found) {
        o->nghbr[4 * ic + in] = jc;
        o->nghbr[4 * jc + im] = ic;
      } else {
        o->nghbr[4 * ic + in] = -1;
      }
    }
  }
  /* nghbre, 372-395: tri locedge (2,3,1), quad (3,4,1,2) (1-based) */
  static const int loc_tri[3] = {1, 2, 0}, loc_quad[4] = {2, 3, 0, 1};
  for (int ic = 0; ic < N; ++ic) {
    int nv = o->nvrt[ic];
    for (int k = 0; k < nv; ++k) {
      int ie = (nv == 3) ? loc_tri[k] : loc_quad[k];
      o->nghbre[4 * ic + ie] = o->nghbr[4 * ic + k];
    }
  }
  /* edge creation, 401-486.  Local edge e=(v_e, v_e+1) checks nghbr index
   * tri (3,1,2), quad (3,4,1,2) (1-based). */
  static const int chk_tri[3] = {2, 0, 1}, chk_quad[4] = {2, 3, 0, 1};
  int ne = 0;
  for (int ic = 0; ic < N; ++ic) {
    int nv = o->nvrt[ic];
    for (int e = 0; e < nv; ++e) {
      int nb = o->nghbr[4 * ic + ((nv == 3) ? chk_tri[e] : chk_quad[e])];
      if (nb > ic || nb == -1) ne++;
    }
  }
  o->nedges = ne;
  o->en1 = xcalloc(ne, sizeof(int)); o->en2 = xcalloc(ne, sizeof(int));
  o->ec1 = xcalloc(ne, sizeof(int)); o->ec2 = xcalloc(ne, sizeof(int));
  o->ex = xcalloc(ne, sizeof(double)); o->ey = xcalloc(ne, sizeof(double));
  o->ea = xcalloc(ne, sizeof(double));
  o->enx = xcalloc(ne, sizeof(double)); o->eny = xcalloc(ne, sizeof(double));
  o->esx = xcalloc(ne, sizeof(double)); o->esy = xcalloc(ne, sizeof(double));
  ne = 0;
  for (int ic = 0; ic < N; ++ic) {
    int nv = o->nvrt[ic];
    for (int e = 0; e < nv; ++e) {
      int nb = o->nghbr[4 * ic + ((nv == 3) ? chk_tri[e] : chk_quad[e])];
      if (nb > ic || nb == -1) {
        o->en1[ne] = o->cnode[4 * ic + e];
        o->en2[ne] = o->cnode[4 * ic + ((e + 1) % nv)];
        o->ec1[ne] = ic;
        o->ec2[ne] = nb;
        ne++;
      }
    }
  }
  /* cell -> edge, 492-624; tri_v2e (2,3,1), qud_v2e (3,4,1,2) */
  static const int v2e_tri[3] = {1, 2, 0}, v2e_quad[4] = {2, 3, 0, 1};
  ne = 0;
  for (int ic = 0; ic < N; ++ic) {
    int nv = o->nvrt[ic];
    for (int e = 0; e < nv; ++e) {
      int jc = o->nghbr[4 * ic + ((nv == 3) ? chk_tri[e] : chk_quad[e])];
      if (jc > ic || jc == -1) {
        o->cedge[4 * ic + e] = ne++;
      } else {
        int nj = o->nvrt[jc];
        for (int k = 0; k < nj; ++k)
          if (o->nghbr[4 * jc + k] == ic) {
            int ie = (nj == 3) ? v2e_tri[k] : v2e_quad[k];
            o->cedge[4 * ic + e] = o->cedge[4 * jc + ie];
            break;
          }
      }
    }
  }
  /* edge geometry, 630-647 */
  for (int i = 0; i < o->nedges; ++i) {
    int v1 = o->en1[i], v2 = o->en2[i];
    double dx = o->nxy[2 * v2] - o->nxy[2 * v1];
    double dy = o->nxy[2 * v2 + 1] - o->nxy[2 * v1 + 1];
    o->ea[i] = sqrt(dx * dx + dy * dy);
    o->ex[i] = 0.5 * (o->nxy[2 * v1] + o->nxy[2 * v2]);
    o->ey[i] = 0.5 * (o->nxy[2 * v1 + 1] + o->nxy[2 * v2 + 1]);
    o->enx[i] = dy / o->ea[i];
    o->eny[i] = -dx / o->ea[i];
    /* builder-new periodic shift: c2's frame -> c1's frame via node n1 */
    if (o->canon && o->ec2[i] >= 0) {
      double px, py;
      cell_node_xy(o, o->ec2[i], o->canon[v1], &px, &py);
      o->esx[i] = o->nxy[2 * v1] - px;
      o->esy[i] = o->nxy[2 * v1 + 1] - py;
    }
  }
  /* nrmlsign, 672-691 */
  for (int ic = 0; ic < N; ++ic)
    for (int ie = 0; ie < o->nvrt[ic]; ++ie) {
      int je = o->cedge[4 * ic + ie];
      o->nrmlsign[4 * ic + ie] = (o->ec1[je] != ic) ? -1.0 : 1.0;
    }
  /* interior / boundary cells, 697-728 */
  o->cell_intr = xcalloc(N, sizeof(int));
  o->ncells_intr = 0; o->ncells_bndr = 0;
  for (int ic = 0; ic < N; ++ic) {
    int im = 1;
    for (int ie = 0; ie < o->nvrt[ic]; ++ie)
      if (o->nghbre[4 * ic + ie] == -1) im = 0;
    if (im) o->cell_intr[o->ncells_intr++] = ic;
    else o->ncells_bndr++;
  }
  int nbc = 0;
  for (int ib = 0; ib < o->nb; ++ib) nbc += o->bnd_ncells[ib];
  if (o->ncells_bndr != nbc) {
    set_err("boundary cells determined %d != listed %d (src/grid_procs.f90:722-728)", o->ncells_bndr, nbc);
    return -1;
  }
  /* interior / boundary edges, 733-791 */
  o->edge_intr = xcalloc(o->nedges, sizeof(int));
  o->nedges_intr = 0; o->nedges_bndr = 0;
  for (int ie = 0; ie < o->nedges; ++ie) {
    if (o->ec1[ie] >= 0 && o->ec2[ie] == -1) o->nedges_bndr++;
    else if (o->ec1[ie] >= 0 && o->ec2[ie] >= 0) o->edge_intr[o->nedges_intr++] = ie;
  }
  o->bnd_nedges = xcalloc(o->nb + 1, sizeof(int));
  o->bnd_eptr = xcalloc(o->nb + 1, sizeof(int));
  o->bnd_edges = xcalloc(o->nedges_bndr + 1, sizeof(int));
  int tot = 0;
  for (int ib = 0; ib < o->nb; ++ib) {
    o->bnd_eptr[ib] = tot;
    for (int i = 0; i < o->bnd_ncells[ib]; ++i) {
      int ic = o->bnd_cells[o->bnd_cptr[ib] + i];
      for (int ie = 0; ie < o->nvrt[ic]; ++ie) {
        int je = o->cedge[4 * ic + ie];
        if (o->ec1[je] == ic && o->ec2[je] == -1) {
          if (tot >= o->nedges_bndr) { set_err("too many boundary edges listed"); return -1; }
          o->bnd_edges[tot++] = je;
          o->bnd_nedges[ib]++;
        }
      }
    }
  }
  o->bnd_eptr[o->nb] = tot;
  if (tot != o->nedges_bndr) {
    set_err("boundary edges determined %d != read %d (src/grid_procs.f90:785-791)", o->nedges_bndr, tot);
    return -1;
  }
  /* src/residual.f90:112-126 pairs bndry%cell(i) with bndry%edge(i); reading
   * past ncells is out of bounds in the reference, so it is an error here */
  for (int ib = 0; ib < o->nb; ++ib)
    if (o->bnd_nedges[ib] > o->bnd_ncells[ib]) {
      set_err("boundary %d has %d edges for %d cells (cell/edge pairing of src/residual.f90:115-125)",
              ib + 1, o->bnd_nedges[ib], o->bnd_ncells[ib]);
      return -1;
    }
  /* c1 > 0 for every edge, 845-852 */
  for (int ie = 0; ie < o->nedges; ++ie)
    if (o->ec1[ie] < 0) { set_err("edge %d has no c1", ie); return -1; }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* LSQ stencils: src/gradient_lsq.f90:70-365                                */
/* ------------------------------------------------------------------------ */
typedef struct { double d2; int idx; } knn_t;
static int knn_cmp(const void *a, const void *b) {
  const knn_t *x = a, *y = b;
  if (x->d2 < y->d2) return -1;
  if (x->d2 > y->d2) return 1;
  return (x->idx > y->idx) - (x->idx < y->idx);
}

/* 8 nearest cell centres, sorted by squared distance (kdtree2_n_nearest,
 * src/kdtree2.f90:1026-1067; includes the query cell itself).  Exact ties are
 * broken by ascending cell index (kdtree2's tie order is unspecified). */
static void nearest8(const orc_t *o, double qx, double qy, int *out) {
  int N = o->ncells;
  knn_t best[9];
  int nbest = 0;
  for (int j = 0; j < N; ++j) {
    double dx = o->cx[j] - qx, dy = o->cy[j] - qy;
    knn_t c = {dx * dx + dy * dy, j};
    if (nbest < 8) {
      best[nbest++] = c;
      qsort(best, nbest, sizeof(knn_t), knn_cmp);
    } else if (knn_cmp(&c, &best[7]) < 0) {
      best[7] = c;
      qsort(best, 8, sizeof(knn_t), knn_cmp);
    }
  }
  for (int i = 0; i < 8; ++i) out[i] = (i < nbest) ? best[i].idx : -1;
}

/* weights and G^-1 D^T coefficients for one cell's stencil (shared by
 * setup_fn 137-201 and setup_nn 277-348; identical arithmetic) */
static void lsq_coefs(orc_t *o, int ic) {
  double xc = o->cx[ic], yc = o->cy[ic];
  int b = o->lsq_ptr[ic], n = o->lsq_ptr[ic + 1] - b;
  double *d1 = xcalloc(n, sizeof(double)), *d2 = xcalloc(n, sizeof(double));
  for (int in = 0; in < n; ++in) {
    int jc = o->lsq_cell[b + in];
    double sx, sy;
    pair_shift(o, ic, jc, &sx, &sy);
    double xc1 = o->cx[jc] + sx, yc1 = o->cy[jc] + sy;
    double dis = sqrt((xc1 - xc) * (xc1 - xc) + (yc1 - yc) * (yc1 - yc));
    double w = 0.0;
    if (dis > 0.0) {
      w = 1.0 / pow(dis, o->op.lsq_p);
      o->lsq_w[b + in] = w;
    }
    d1[in] = w * (xc1 - xc);
    d2[in] = w * (yc1 - yc);
  }
  double g11 = 0, g12 = 0, g21 = 0, g22 = 0;
  for (int in = 0; in < n; ++in) g11 = g11 + d1[in] * d1[in];
  for (int in = 0; in < n; ++in) g12 = g12 + d1[in] * d2[in];
  for (int in = 0; in < n; ++in) g21 = g21 + d2[in] * d1[in];
  for (int in = 0; in < n; ++in) g22 = g22 + d2[in] * d2[in];
  double det = g11 * g22 - g12 * g21;
  double gi11 = 1.0 / det * g22, gi22 = 1.0 / det * g11;
  double gi12 = -(1.0 / det * g12), gi21 = -(1.0 / det * g21);
  for (int in = 0; in < n; ++in) {
    double c1 = 0.0, c2 = 0.0;
    c1 = c1 + gi11 * d1[in];
    c1 = c1 + gi12 * d2[in];
    c2 = c2 + gi21 * d1[in];
    c2 = c2 + gi22 * d2[in];
    o->lsq_coef[2 * (b + in)] = c1;
    o->lsq_coef[2 * (b + in) + 1] = c2;
  }
  free(d1); free(d2);
}

static int setup_lsq(orc_t *o) {
  int N = o->ncells;
  o->lsq_ptr = xcalloc(N + 1, sizeof(int));
  if (!o->op.lsq_nn) {
    /* setup_fn, 88-131: stencil = cell%nghbr (vertex-opposite order) with
     * boundary gaps filled from the 8 nearest centres */
    for (int ic = 0; ic < N; ++ic) o->lsq_ptr[ic + 1] = o->lsq_ptr[ic] + o->nvrt[ic];
    o->lsq_cell = xcalloc(o->lsq_ptr[N], sizeof(int));
    for (int ic = 0; ic < N; ++ic) {
      int b = o->lsq_ptr[ic], nv = o->nvrt[ic];
      int tmpi[4] = {0, 0, 0, 0}, izb = 0;
      for (int in = 0; in < nv; ++in) {
        int jc = o->nghbr[4 * ic + in];
        o->lsq_cell[b + in] = -1;
        if (jc == -1) tmpi[izb++] = in;
        else o->lsq_cell[b + in] = jc;
      }
      if (izb > 0 && izb < 3) {
        int cand[8];
        nearest8(o, o->cx[ic], o->cy[ic], cand);
        int nt = 0;
        for (int i = 0; i < 8; ++i) {
          int jc = cand[i];
          if (jc < 0 || jc == ic) continue;
          int isnb = 0;
          for (int in = 0; in < nv; ++in)
            if (o->nghbr[4 * ic + in] == jc) isnb = 1;
          if (!isnb) o->lsq_cell[b + tmpi[nt++]] = jc;
          if (nt == izb) break;
        }
        if (nt < izb) { set_err("setup_fn: not enough augmentation cells for cell %d", ic); return -1; }
      } else if (izb > 2) {
        set_err("setup_fn: #s of edges on the boundary>2 at cell %d (src/gradient_lsq.f90:126-128)", ic);
        return -1;
      }
    }
  } else {
    /* setup_nn, 227-271: all cells sharing a vertex, ascending, unique */
    int *tmp = xcalloc(64, sizeof(int));
    int cap = 64;
    int **lists = xcalloc(N, sizeof(int *));
    int *cnt = xcalloc(N, sizeof(int));
    for (int ic = 0; ic < N; ++ic) {
      int nt = 0;
      for (int in = 0; in < o->nvrt[ic]; ++in) {
        int iv = o->cnodec[4 * ic + in];
        for (int i = o->n2c_ptr[iv]; i < o->n2c_ptr[iv + 1]; ++i) {
          int jc = o->n2c[i];
          if (jc != ic) {
            if (nt == cap) { cap *= 2; tmp = realloc(tmp, cap * sizeof(int)); }
            tmp[nt++] = jc;
          }
        }
      }
      /* repeated min over values greater than the previous min (253-261) */
      int *unq = xcalloc(nt + 1, sizeof(int));
      int nu = 0;
      long cur = -1;
      for (;;) {
        long m = -1;
        for (int i = 0; i < nt; ++i)
          if (tmp[i] > cur && (m < 0 || tmp[i] < m)) m = tmp[i];
        if (m < 0) break;
        unq[nu++] = (int)m;
        cur = m;
      }
      lists[ic] = unq;
      cnt[ic] = nu;
    }
    for (int ic = 0; ic < N; ++ic) o->lsq_ptr[ic + 1] = o->lsq_ptr[ic] + cnt[ic];
    o->lsq_cell = xcalloc(o->lsq_ptr[N], sizeof(int));
    for (int ic = 0; ic < N; ++ic) {
      memcpy(&o->lsq_cell[o->lsq_ptr[ic]], lists[ic], cnt[ic] * sizeof(int));
      free(lists[ic]);
    }
    free(lists); free(cnt); free(tmp);
  }
  o->lsq_w = xcalloc(o->lsq_ptr[N], sizeof(double));
  o->lsq_coef = xcalloc(2 * o->lsq_ptr[N], sizeof(double));
  for (int ic = 0; ic < N; ++ic) lsq_coefs(o, ic);
  /* grad_lsq_verify, 490-529: exact for f = 2x + y to 1e-10 */
  for (int ic = 0; ic < N; ++ic) {
    double dfx = 0.0, dfy = 0.0;
    for (int in = o->lsq_ptr[ic]; in < o->lsq_ptr[ic + 1]; ++in) {
      int jc = o->lsq_cell[in];
      double sx, sy;
      pair_shift(o, ic, jc, &sx, &sy);
      double xnb = o->cx[jc] + sx, ynb = o->cy[jc] + sy;
      double diff = 1.0 * ynb + 2.0 * xnb - (1.0 * o->cy[ic] + 2.0 * o->cx[ic]);
      dfx = dfx + o->lsq_coef[2 * in] * diff * o->lsq_w[in];
      dfy = dfy + o->lsq_coef[2 * in + 1] * diff * o->lsq_w[in];
    }
    if (fabs(dfx - 2.0) > 1e-10 || fabs(dfy - 1.0) > 1e-10) {
      set_err("LSQ coefficients are not correct at cell %d (src/gradient_lsq.f90:513-524)", ic);
      return -1;
    }
  }
  return 0;
}

/* centroid of the cell on side `c` of edge je, in the edge's (c1) frame */
static void side_centroid(const orc_t *o, int je, int c, double *x, double *y) {
  if (c == o->ec1[je]) { *x = o->cx[c]; *y = o->cy[c]; }
  else { *x = o->cx[c] + o->esx[je]; *y = o->cy[c] + o->esy[je]; }
}

/* GGCB setup: src/gradient_ggcb.f90:48-110 */
static void setup_ggcb(orc_t *o) {
  int N = o->ncells;
  o->ggcb_ptr = xcalloc(4 * N, sizeof(int));
  o->ggcb_coef0 = xcalloc(2 * N, sizeof(double));
  o->ggcb_coefnb = xcalloc(8 * N, sizeof(double));
  for (int ic = 0; ic < N; ++ic) {
    for (int ie = 0; ie < o->nvrt[ic]; ++ie) {
      int je = o->cedge[4 * ic + ie];
      int ic1 = o->ec1[je], ic2 = o->ec2[je];
      double af = o->ea[je];
      double nxf = o->enx[je] * o->nrmlsign[4 * ic + ie];
      double nyf = o->eny[je] * o->nrmlsign[4 * ic + ie];
      double xf = o->ex[je], yf = o->ey[je];
      double xc, yc;
      side_centroid(o, je, ic, &xc, &yc);
      double dx = xf - xc, dy = yf - yc;
      double d0 = sqrt(dx * dx + dy * dy);
      double xc1 = xc, yc1 = yc;
      o->ggcb_ptr[4 * ic + ie] = ic;
      if (ic1 >= 0 && ic1 != ic) {
        side_centroid(o, je, ic1, &xc1, &yc1);
        o->ggcb_ptr[4 * ic + ie] = ic1;
      } else if (ic2 >= 0 && ic2 != ic) {
        side_centroid(o, je, ic2, &xc1, &yc1);
        o->ggcb_ptr[4 * ic + ie] = ic2;
      }
      dx = xf - xc1; dy = yf - yc1;
      double d1 = sqrt(dx * dx + dy * dy);
#if ORC_MUTANT == MUT_GGCB_SWAP_WEIGHTS
      { double tmp = d0; d0 = d1; d1 = tmp; }
#endif
      o->ggcb_coef0[2 * ic] = o->ggcb_coef0[2 * ic] + d1 / (d0 + d1) * nxf * af;
      o->ggcb_coef0[2 * ic + 1] = o->ggcb_coef0[2 * ic + 1] + d1 / (d0 + d1) * nyf * af;
      o->ggcb_coefnb[8 * ic + 2 * ie] = d0 / (d0 + d1) * nxf * af;
      o->ggcb_coefnb[8 * ic + 2 * ie + 1] = d0 / (d0 + d1) * nyf * af;
    }
  }
}

/* GGNB setup: src/gradient_ggnb.f90:217-345 */
static void setup_ggnb(orc_t *o) {
  int N = o->ncells, NN = o->nnodes;
  double *idw = xcalloc(NN, sizeof(double));
  for (int v = 0; v < NN; ++v) {
    if (o->n2c_ptr[v + 1] == o->n2c_ptr[v]) continue;
    double idt = 0.0;
    for (int i = o->n2c_ptr[v]; i < o->n2c_ptr[v + 1]; ++i) {
      int ic = o->n2c[i];
      double xv, yv;
      cell_node_xy(o, ic, v, &xv, &yv);
      double dx = o->cx[ic] - xv, dy = o->cy[ic] - yv;
      double d = sqrt(dx * dx + dy * dy);
#if ORC_MUTANT == MUT_GGNB_IDW_SQUARE
      d = d * d;
#endif
      idt = idt + 1.0 / d;
    }
    idw[v] = 1.0 / idt;
  }
  o->ggnb_coef0 = xcalloc(2 * N, sizeof(double));
  for (int ic = 0; ic < N; ++ic) {
    for (int ie = 0; ie < o->nvrt[ic]; ++ie) {
      int je = o->cedge[4 * ic + ie];
      double af = o->ea[je];
      double nxf = o->enx[je] * o->nrmlsign[4 * ic + ie];
      double nyf = o->eny[je] * o->nrmlsign[4 * ic + ie];
      int iv1 = o->canon ? o->canon[o->en1[je]] : o->en1[je];
      int iv2 = o->canon ? o->canon[o->en2[je]] : o->en2[je];
      double xv, yv;
      cell_node_xy(o, ic, iv1, &xv, &yv);
      double dx = o->cx[ic] - xv, dy = o->cy[ic] - yv;
      double w1 = 1.0 / sqrt(dx * dx + dy * dy);
      cell_node_xy(o, ic, iv2, &xv, &yv);
      dx = o->cx[ic] - xv; dy = o->cy[ic] - yv;
      double w2 = 1.0 / sqrt(dx * dx + dy * dy);
#if ORC_MUTANT == MUT_GGNB_IDW_SQUARE
      w1 = w1 * w1;
      w2 = w2 * w2;
#endif
      double wt1 = idw[iv1], wt2 = idw[iv2];
      o->ggnb_coef0[2 * ic] = o->ggnb_coef0[2 * ic] + af * nxf / 2.0 * (w1 * wt1 + w2 * wt2);
      o->ggnb_coef0[2 * ic + 1] = o->ggnb_coef0[2 * ic + 1] + af * nyf / 2.0 * (w1 * wt1 + w2 * wt2);
    }
  }
  o->ggnb_ptr = xcalloc(N + 1, sizeof(int));
  for (int ic = 0; ic < N; ++ic) {
    int i = 0;
    for (int in = 0; in < o->nvrt[ic]; ++in) {
      int iv = o->cnodec[4 * ic + in];
      i += o->n2c_ptr[iv + 1] - o->n2c_ptr[iv];
    }
    o->ggnb_ptr[ic + 1] = o->ggnb_ptr[ic] + i;
  }
  int nnz = o->ggnb_ptr[N];
  o->ggnb_cell = xcalloc(nnz, sizeof(int));
  o->ggnb_in = xcalloc(nnz, sizeof(int));
  o->ggnb_coefnb = xcalloc(nnz, sizeof(double));
  for (int ic = 0; ic < N; ++ic) {
    int i = o->ggnb_ptr[ic];
    for (int in = 0; in < o->nvrt[ic]; ++in) {
      int iv = o->cnodec[4 * ic + in];
      for (int j = o->n2c_ptr[iv]; j < o->n2c_ptr[iv + 1]; ++j) {
        int jc = o->n2c[j];
        double xv, yv;
        cell_node_xy(o, jc, iv, &xv, &yv);
        double dx = o->cx[jc] - xv, dy = o->cy[jc] - yv;
        double d = sqrt(dx * dx + dy * dy);
#if ORC_MUTANT == MUT_GGNB_IDW_SQUARE
        d = d * d;
#endif
        o->ggnb_cell[i] = jc;
        o->ggnb_in[i] = in;
        o->ggnb_coefnb[i] = idw[iv] / d;
#if ORC_MUTANT != MUT_GGNB_SELF_TWICE
        if (jc == ic) o->ggnb_coefnb[i] = 0.0;
#endif
        i++;
      }
    }
  }
  o->ggnb_coefedg = xcalloc(8 * N, sizeof(double));
  for (int ic = 0; ic < N; ++ic) {
    for (int in = 0; in < o->nvrt[ic]; ++in) {
      int iv = o->cnodec[4 * ic + in];
      for (int ie = 0; ie < o->nvrt[ic]; ++ie) {
        int je = o->cedge[4 * ic + ie];
        for (int j = 0; j < 2; ++j) {
          int jv = (j == 0) ? o->en1[je] : o->en2[je];
          if (o->canon) jv = o->canon[jv];
          if (jv == iv) {
            double af = o->ea[je];
            double nxf = o->enx[je] * o->nrmlsign[4 * ic + ie];
            double nyf = o->eny[je] * o->nrmlsign[4 * ic + ie];
            o->ggnb_coefedg[8 * ic + 2 * in] = o->ggnb_coefedg[8 * ic + 2 * in] + af * nxf / 2.0;
            o->ggnb_coefedg[8 * ic + 2 * in + 1] = o->ggnb_coefedg[8 * ic + 2 * in + 1] + af * nyf / 2.0;
          }
        }
      }
    }
  }
  free(idw);
}

/* ------------------------------------------------------------------------ */
orc_t *orc_create(int nnodes, const double *xy, int ntri, int nquad, const int *cell_nodes,
                  const int *node_canon, int nb, const int *bnd_ncells, const int *bnd_types,
                  const int *bnd_cells, const orc_opts *opts) {
  g_err[0] = 0;
  orc_t *o = xcalloc(1, sizeof(orc_t));
  o->op = *opts;
  /* src/input.f90:258-271: kappa forced to 0 for upwind-1st/2nd */
  o->kappa = (opts->recon == 3) ? opts->kappa : 0.0;
  o->nnodes = nnodes; o->ntri = ntri; o->nquad = nquad;
  int N = o->ncells = ntri + nquad;
  o->nxy = xcalloc(2 * nnodes, sizeof(double));
  memcpy(o->nxy, xy, 2 * nnodes * sizeof(double));
  if (node_canon) {
    o->canon = xcalloc(nnodes, sizeof(int));
    memcpy(o->canon, node_canon, nnodes * sizeof(int));
  }
  o->nvrt = xcalloc(N, sizeof(int));
  o->cnode = xcalloc(4 * N, sizeof(int));
  o->cnodec = xcalloc(4 * N, sizeof(int));
  o->nghbr = xcalloc(4 * N, sizeof(int));
  o->nghbre = xcalloc(4 * N, sizeof(int));
  o->cedge = xcalloc(4 * N, sizeof(int));
  o->nrmlsign = xcalloc(4 * N, sizeof(double));
  o->cx = xcalloc(N, sizeof(double)); o->cy = xcalloc(N, sizeof(double));
  o->vol = xcalloc(N, sizeof(double));
  const int *p = cell_nodes;
  for (int ic = 0; ic < N; ++ic) {
    int nv = (ic < ntri) ? 3 : 4;
    o->nvrt[ic] = nv;
    for (int k = 0; k < 4; ++k) o->cnode[4 * ic + k] = -1, o->cnodec[4 * ic + k] = -1, o->nghbr[4 * ic + k] = -1,
                                o->nghbre[4 * ic + k] = -1, o->cedge[4 * ic + k] = -1;
    for (int k = 0; k < nv; ++k) {
      int v = *p++;
      if (v < 0 || v >= nnodes) { set_err("cell %d references node %d", ic, v); orc_destroy(o); return NULL; }
      o->cnode[4 * ic + k] = v;
      o->cnodec[4 * ic + k] = node_canon ? node_canon[v] : v;
    }
  }
  o->nb = nb;
  o->bnd_type = xcalloc(nb + 1, sizeof(int));
  o->bnd_ncells = xcalloc(nb + 1, sizeof(int));
  o->bnd_cptr = xcalloc(nb + 1, sizeof(int));
  int tot = 0;
  for (int ib = 0; ib < nb; ++ib) {
    o->bnd_type[ib] = bnd_types[ib];
    o->bnd_ncells[ib] = bnd_ncells[ib];
    o->bnd_cptr[ib] = tot;
    tot += bnd_ncells[ib];
  }
  o->bnd_cptr[nb] = tot;
  o->bnd_cells = xcalloc(tot + 1, sizeof(int));
  if (tot) memcpy(o->bnd_cells, bnd_cells, tot * sizeof(int));
  if (grid_data(o)) { orc_destroy(o); return NULL; }
  /* gradient_init: src/gradient.f90:18-38.  The limiter always reads the LSQ
   * stencil (src/gradient_limiter.f90:54-58); for GG gradients the node-
   * neighbour LSQ stencil is built for it (builder choice). */
  if (opts->grad_method == 3 || opts->limiter > 0) {
    orc_opts save = o->op;
    if (opts->grad_method != 3) o->op.lsq_nn = 1;
    if (setup_lsq(o)) { orc_destroy(o); return NULL; }
    o->op = save;
  }
  if (opts->grad_method == 1) setup_ggcb(o);
  if (opts->grad_method == 2) setup_ggnb(o);
  o->pvar = xcalloc(4 * N, sizeof(double));
  o->cvar = xcalloc(4 * N, sizeof(double));
  o->grad = xcalloc(8 * N, sizeof(double));
  o->resid = xcalloc(4 * N, sizeof(double));
  o->phi = xcalloc(N, sizeof(double));
  o->ws = xcalloc(N, sizeof(double));
  o->dt_local = xcalloc(N, sizeof(double));
  for (int i = 0; i < N; ++i) o->dt_local[i] = opts->dt;
  /* data_solution_init: src/data_solution.f90:41-66 */
  o->pvar_inf[0] = 1.0;
  o->pvar_inf[1] = opts->mach * cosd(opts->aoa_deg);
  o->pvar_inf[2] = opts->mach * sind(opts->aoa_deg);
  o->pvar_inf[3] = 1.0 / opts->gamma;
  return o;
}

void orc_destroy(orc_t *o) {
  if (!o) return;
  void *ptrs[] = {o->nxy, o->canon, o->nvrt, o->cnode, o->cnodec, o->nghbr, o->nghbre, o->cedge, o->cx, o->cy,
                  o->vol, o->nrmlsign, o->n2c_ptr, o->n2c, o->en1, o->en2, o->ec1, o->ec2, o->ex, o->ey, o->ea,
                  o->enx, o->eny, o->esx, o->esy, o->edge_intr, o->cell_intr, o->bnd_type, o->bnd_ncells,
                  o->bnd_cptr, o->bnd_cells, o->bnd_nedges, o->bnd_eptr, o->bnd_edges, o->lsq_ptr, o->lsq_cell,
                  o->lsq_w, o->lsq_coef, o->ggcb_ptr, o->ggcb_coef0, o->ggcb_coefnb, o->ggnb_coef0,
                  o->ggnb_coefedg, o->ggnb_ptr, o->ggnb_cell, o->ggnb_in, o->ggnb_coefnb, o->pvar, o->cvar,
                  o->grad, o->resid, o->phi, o->ws, o->dt_local, o->cf_ptr, o->cf, o->fa, o->wa};
  for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; ++i) free(ptrs[i]);
  free(o);
}

void orc_sizes(const orc_t *o, int *s) {
  s[0] = o->ncells; s[1] = o->nedges; s[2] = o->nnodes; s[3] = o->nedges_intr;
  s[4] = o->nedges_bndr; s[5] = o->ncells_intr; s[6] = o->ncells_bndr; s[7] = o->nb;
}
void orc_get_edges(const orc_t *o, int *n1, int *n2, int *c1, int *c2, double *g) {
  int E = o->nedges;
  memcpy(n1, o->en1, E * sizeof(int)); memcpy(n2, o->en2, E * sizeof(int));
  memcpy(c1, o->ec1, E * sizeof(int)); memcpy(c2, o->ec2, E * sizeof(int));
  memcpy(g, o->ex, E * 8); memcpy(g + E, o->ey, E * 8); memcpy(g + 2 * E, o->ea, E * 8);
  memcpy(g + 3 * E, o->enx, E * 8); memcpy(g + 4 * E, o->eny, E * 8);
  memcpy(g + 5 * E, o->esx, E * 8); memcpy(g + 6 * E, o->esy, E * 8);
}
void orc_get_cells(const orc_t *o, double *xyv, int *nvrt, int *nghbre, int *edge, double *nrmlsign) {
  int N = o->ncells;
  memcpy(xyv, o->cx, N * 8); memcpy(xyv + N, o->cy, N * 8); memcpy(xyv + 2 * N, o->vol, N * 8);
  memcpy(nvrt, o->nvrt, N * sizeof(int));
  memcpy(nghbre, o->nghbre, 4 * N * sizeof(int));
  memcpy(edge, o->cedge, 4 * N * sizeof(int));
  memcpy(nrmlsign, o->nrmlsign, 4 * N * 8);
}
void orc_get_lists(const orc_t *o, int *edge_intr, int *cell_intr, int *bnd_nedges, int *bnd_edges) {
  memcpy(edge_intr, o->edge_intr, o->nedges_intr * sizeof(int));
  memcpy(cell_intr, o->cell_intr, o->ncells_intr * sizeof(int));
  memcpy(bnd_nedges, o->bnd_nedges, o->nb * sizeof(int));
  memcpy(bnd_edges, o->bnd_edges, o->nedges_bndr * sizeof(int));
}
int orc_lsq_nnz(const orc_t *o) { return o->lsq_ptr ? o->lsq_ptr[o->ncells] : 0; }
void orc_get_lsq(const orc_t *o, int *ptr, int *cell, double *w, double *coef) {
  int N = o->ncells, nnz = orc_lsq_nnz(o);
  if (!o->lsq_ptr) return;
  memcpy(ptr, o->lsq_ptr, (N + 1) * sizeof(int));
  memcpy(cell, o->lsq_cell, nnz * sizeof(int));
  memcpy(w, o->lsq_w, nnz * 8);
  memcpy(coef, o->lsq_coef, 2 * nnz * 8);
}

void orc_set_cvar(orc_t *o, const double *c) { memcpy(o->cvar, c, 4 * o->ncells * 8); }
void orc_get_cvar(const orc_t *o, double *c) { memcpy(c, o->cvar, 4 * o->ncells * 8); }
void orc_get_pvar(const orc_t *o, double *c) { memcpy(c, o->pvar, 4 * o->ncells * 8); }
void orc_set_pvar(orc_t *o, const double *c) { memcpy(o->pvar, c, 4 * o->ncells * 8); }
void orc_get_grad(const orc_t *o, double *g) { memcpy(g, o->grad, 8 * o->ncells * 8); }
void orc_get_phi(const orc_t *o, double *g) { memcpy(g, o->phi, o->ncells * 8); }
void orc_get_resid(const orc_t *o, double *g) { memcpy(g, o->resid, 4 * o->ncells * 8); }
void orc_get_ws(const orc_t *o, double *g) { memcpy(g, o->ws, o->ncells * 8); }
void orc_get_dt_local(const orc_t *o, double *g) { memcpy(g, o->dt_local, o->ncells * 8); }

/* ------------------------------------------------------------------------ */
/* data_solution: src/data_solution.f90:72-106                              */
/* ------------------------------------------------------------------------ */
void orc_cvar2pvar(orc_t *o) {
  double g = o->op.gamma;
  ORC_PAR
  for (int ic = 0; ic < o->ncells; ++ic) {
    double *pv = &o->pvar[4 * ic];
    const double *cv = &o->cvar[4 * ic];
    pv[0] = cv[0];
    pv[1] = cv[1] / cv[0];
    pv[2] = cv[2] / cv[0];
    pv[3] = (g - 1.0) * (cv[3] - 0.5 * pv[0] * (pv[1] * pv[1] + pv[2] * pv[2]));
  }
}
void orc_pvar2cvar(orc_t *o) {
  double g = o->op.gamma;
  for (int ic = 0; ic < o->ncells; ++ic) {
    const double *pv = &o->pvar[4 * ic];
    double *cv = &o->cvar[4 * ic];
    cv[0] = pv[0];
    cv[1] = pv[0] * pv[1];
    cv[2] = pv[0] * pv[2];
    cv[3] = pv[3] / (g - 1.0) + 0.5 * pv[0] * (pv[1] * pv[1] + pv[2] * pv[2]);
  }
}

/* ------------------------------------------------------------------------ */
/* gradients: src/gradient.f90:44-96, src/gradient_lsq.f90:371-441,         */
/* src/gradient_ggcb.f90:116-166, src/gradient_ggnb.f90:351-411             */
/* ------------------------------------------------------------------------ */
static void grad_generic(const orc_t *o, int nvar, const double *q, int qstride, double *gx, double *gy,
                         int gstride) {
  int N = o->ncells;
  int m = o->op.grad_method;
  if (m == 3) {
    ORC_PAR
    for (int ic = 0; ic < N; ++ic) {
      for (int k = 0; k < nvar; ++k) gx[gstride * ic + k] = 0.0, gy[gstride * ic + k] = 0.0;
      for (int in = o->lsq_ptr[ic]; in < o->lsq_ptr[ic + 1]; ++in) {
        int jc = o->lsq_cell[in];
        for (int iv = 0; iv < nvar; ++iv) {
          double diff = q[qstride * jc + iv] - q[qstride * ic + iv];
          gx[gstride * ic + iv] = gx[gstride * ic + iv] + o->lsq_coef[2 * in] * diff * o->lsq_w[in];
          gy[gstride * ic + iv] = gy[gstride * ic + iv] + o->lsq_coef[2 * in + 1] * diff * o->lsq_w[in];
        }
      }
    }
  } else if (m == 1) {
    for (int iv = 0; iv < nvar; ++iv) {
      ORC_PAR
      for (int ic = 0; ic < N; ++ic) {
        double a = o->ggcb_coef0[2 * ic] * q[qstride * ic + iv];
        double b = o->ggcb_coef0[2 * ic + 1] * q[qstride * ic + iv];
        for (int ie = 0; ie < o->nvrt[ic]; ++ie) {
          int ic1 = o->ggcb_ptr[4 * ic + ie];
          a = a + o->ggcb_coefnb[8 * ic + 2 * ie] * q[qstride * ic1 + iv];
          b = b + o->ggcb_coefnb[8 * ic + 2 * ie + 1] * q[qstride * ic1 + iv];
        }
        gx[gstride * ic + iv] = a / o->vol[ic];
        gy[gstride * ic + iv] = b / o->vol[ic];
      }
    }
  } else if (m == 2) {
    for (int iv = 0; iv < nvar; ++iv) {
      ORC_PAR
      for (int ic = 0; ic < N; ++ic) {
        double a = o->ggnb_coef0[2 * ic] * q[qstride * ic + iv];
        double b = o->ggnb_coef0[2 * ic + 1] * q[qstride * ic + iv];
        for (int i = o->ggnb_ptr[ic]; i < o->ggnb_ptr[ic + 1]; ++i) {
          int jc = o->ggnb_cell[i], in = o->ggnb_in[i];
          a = a + o->ggnb_coefedg[8 * ic + 2 * in] * o->ggnb_coefnb[i] * q[qstride * jc + iv];
          b = b + o->ggnb_coefedg[8 * ic + 2 * in + 1] * o->ggnb_coefnb[i] * q[qstride * jc + iv];
        }
        gx[gstride * ic + iv] = a / o->vol[ic];
        gy[gstride * ic + iv] = b / o->vol[ic];
      }
    }
  }
}

void orc_gradient(orc_t *o) {
  int N = o->ncells;
  memset(o->grad, 0, 8 * N * sizeof(double));
  if (o->op.recon == 1) return; /* src/gradient.f90:49 */
  grad_generic(o, 4, o->pvar, 4, o->grad, o->grad + 4 * N, 4);
}

void orc_gradient_1var(const orc_t *o, const double *fc, double *dfc) {
  int N = o->ncells;
  memset(dfc, 0, 2 * N * sizeof(double));
  grad_generic(o, 1, fc, 1, dfc, dfc + N, 1);
}

/* ------------------------------------------------------------------------ */
/* limiter: src/gradient_limiter.f90:19-134                                 */
/* ------------------------------------------------------------------------ */
double orc_limiter_fn(int lt, double a, double b, double vol) {
  double lim = 0.0;
  if (lt == 1) { /* 'venk' 113-117 */
    double ap = 5.0;
    double h = 2.0 * sqrt(vol / PI);
    double t = ap * h;
    double eps2 = t * t * t;
    lim = ((a * a + eps2) + 2.0 * b * a) / (a * a + 2.0 * (b * b) + a * b + eps2);
  } else if (lt == 2) { /* 'barth' 119-120 */
    double r = a / b;
    lim = (1.0 < r) ? 1.0 : r;
  } else if (lt == 3) { /* 'albada' 122-127 */
    double h = 2.0 * sqrt(vol / PI);
    double t = 0.3 * h;
    double eps2 = t * t * t;
    lim = ((b * b + eps2) * a + (a * a + eps2) * b) / (a * a + b * b + 2.0 * eps2);
    lim = lim / (b + eps2);
  }
  return lim;
}

void orc_limiter(orc_t *o) {
  int N = o->ncells;
  if (o->op.recon == 1) { for (int i = 0; i < N; ++i) o->phi[i] = 0.0; return; }
  if (o->op.limiter == 0) { for (int i = 0; i < N; ++i) o->phi[i] = 1.0; return; }
  double *phi4 = xcalloc(4 * N, sizeof(double));
  const double *gx = o->grad, *gy = o->grad + 4 * N;
  for (int iv = 0; iv < 4; ++iv) {
    ORC_PAR
    for (int ic = 0; ic < N; ++ic) {
      double q = o->pvar[4 * ic + iv];
      double pmin = q, pmax = q;
#if ORC_MUTANT == MUT_LIM_FACE_STENCIL
      for (int ie = 0; ie < o->nvrt[ic]; ++ie) {
        int je = o->cedge[4 * ic + ie];
        int jc = (o->ec1[je] == ic) ? o->ec2[je] : o->ec1[je];
        if (jc < 0) continue;
        double qj = o->pvar[4 * jc + iv];
        pmin = (qj < pmin) ? qj : pmin;
        pmax = (qj > pmax) ? qj : pmax;
      }
#else
      for (int i = o->lsq_ptr[ic]; i < o->lsq_ptr[ic + 1]; ++i) {
        double qj = o->pvar[4 * o->lsq_cell[i] + iv];
        pmin = (qj < pmin) ? qj : pmin;
        pmax = (qj > pmax) ? qj : pmax;
      }
#endif
      double pe[4] = {100.0, 100.0, 100.0, 100.0};
      for (int ie = 0; ie < o->nvrt[ic]; ++ie) {
        int je = o->cedge[4 * ic + ie];
        double xc, yc;
        side_centroid(o, je, ic, &xc, &yc);
#if ORC_MUTANT == MUT_LIM_GRAD_SWAP
        double pf = q + (o->ex[je] - xc) * gy[4 * ic + iv] + (o->ey[je] - yc) * gx[4 * ic + iv];
#else
        double pf = q + (o->ex[je] - xc) * gx[4 * ic + iv] + (o->ey[je] - yc) * gy[4 * ic + iv];
#endif
        double dmax = pmax - q, dmin = pmin - q, diff = pf - q;
        if (diff > 0.0) pe[ie] = orc_limiter_fn(o->op.limiter, dmax, diff, o->vol[ic]);
        else if (diff < 0.0) pe[ie] = orc_limiter_fn(o->op.limiter, dmin, diff, o->vol[ic]);
        else pe[ie] = 1.0;
      }
      double m = pe[0];
      for (int k = 1; k < 4; ++k) m = (pe[k] < m) ? pe[k] : m;
      phi4[4 * ic + iv] = (1.0 < m) ? 1.0 : m;
    }
  }
  ORC_PAR
  for (int ic = 0; ic < N; ++ic) {
    double m = phi4[4 * ic];
#if ORC_MUTANT != MUT_LIM_RHO_ONLY
    for (int k = 1; k < 4; ++k) m = (phi4[4 * ic + k] < m) ? phi4[4 * ic + k] : m;
#endif
    o->phi[ic] = m;
  }
  free(phi4);
}

/* ------------------------------------------------------------------------ */
/* Roe flux: src/flux_invscid.f90:37-136                                    */
/* ------------------------------------------------------------------------ */
void orc_roe(double gamma, const double *pL, const double *pR, double nx, double ny, double *flux,
             double *ws_max) {
  double tx = -ny, ty = nx;
  double rhoL = pL[0], uL = pL[1], vL = pL[2], PL = pL[3];
  double rhoR = pR[0], uR = pR[1], vR = pR[2], PR = pR[3];
  double unL = uL * nx + vL * ny, unR = uR * nx + vR * ny;
  double utL = uL * tx + vL * ty, utR = uR * tx + vR * ty;
  double aL = sqrt(gamma * PL / rhoL), aR = sqrt(gamma * PR / rhoR);
  double kL = 0.5 * (uL * uL + vL * vL), kR = 0.5 * (uR * uR + vR * vR);
  double HL = aL * aL / (gamma - 1.0) + kL, HR = aR * aR / (gamma - 1.0) + kR;
  double RT = sqrt(rhoR / rhoL);
  double rho = RT * rhoL;
  double u = (uL + RT * uR) / (1.0 + RT);
  double v = (vL + RT * vR) / (1.0 + RT);
  double H = (HL + RT * HR) / (1.0 + RT);
  double a = sqrt((gamma - 1.0) * (H - 0.5 * (u * u + v * v)));
  double un = u * nx + v * ny;
  double ut = u * tx + v * ty;
  double drho = rhoR - rhoL, dp = PR - PL, dun = unR - unL, dut = utR - utL;
  double LdU[4], ws[4], Rv[4][4];
  LdU[0] = (dp - rho * a * dun) / (2.0 * a * a);
  LdU[1] = rho * dut;
  LdU[2] = drho - dp / (a * a);
  LdU[3] = (dp + rho * a * dun) / (2.0 * a * a);
  ws[0] = fabs(un - a);
  ws[1] = fabs(un);
  ws[2] = fabs(un);
  ws[3] = fabs(un + a);
  double dws1 = 1.0 / 5.0, dws4 = 1.0 / 5.0;
  if (ws[0] < dws1) ws[0] = 0.5 * (ws[0] * ws[0] / dws1 + dws1);
  if (ws[3] < dws4) ws[3] = 0.5 * (ws[3] * ws[3] / dws4 + dws4);
  double tke = 0.5 * (u * u + v * v);
  Rv[0][0] = 1.0;         Rv[0][1] = 0.0; Rv[0][2] = 1.0; Rv[0][3] = 1.0;
  Rv[1][0] = u - a * nx;  Rv[1][1] = tx;  Rv[1][2] = u;   Rv[1][3] = u + a * nx;
  Rv[2][0] = v - a * ny;  Rv[2][1] = ty;  Rv[2][2] = v;   Rv[2][3] = v + a * ny;
  Rv[3][0] = H - un * a;  Rv[3][1] = ut;  Rv[3][2] = tke; Rv[3][3] = H + un * a;
  double diss[4] = {0, 0, 0, 0};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) diss[i] = diss[i] + ws[j] * LdU[j] * Rv[i][j];
  double fL[4], fR[4];
  fL[0] = rhoL * unL;
  fL[1] = rhoL * unL * uL + PL * nx;
  fL[2] = rhoL * unL * vL + PL * ny;
  fL[3] = rhoL * unL * HL;
  fR[0] = rhoR * unR;
  fR[1] = rhoR * unR * uR + PR * nx;
  fR[2] = rhoR * unR * vR + PR * ny;
  fR[3] = rhoR * unR * HR;
  for (int i = 0; i < 4; ++i) flux[i] = 0.5 * (fL[i] + fR[i] - 1.0 * diss[i]);
  *ws_max = 0.5 * (fabs(un) + a);
}

/* ------------------------------------------------------------------------ */
/* mms: src/mms.f90:124-265                                                 */
/* ------------------------------------------------------------------------ */
static double manufactured_sol(double a0, double as, double ax, double ay, int nx, int ny, double x, double y) {
  double fval;
  if (nx + ny == 0) {
    fval = a0 + as * sin(ax * x + ay * y);
  } else if ((nx + ny) % 2 == 0) {
    fval = -(pow(ax, nx) * pow(ay, ny)) * as * sin(ax * x + ay * y);
    if ((nx + ny) % 4 == 0) fval = -fval;
  } else {
    fval = (pow(ax, nx) * pow(ay, ny)) * as * cos(ax * x + ay * y);
    if ((nx + ny + 1) % 4 == 0) fval = -fval;
  }
  return fval;
}

void orc_mms_state(const orc_t *o, double xc, double yc, double *sol, double *rhs) {
  /* constants, src/mms.f90:79-101 */
  const double cr0 = 1.12, crs = 0.15, crx = 3.12 * PI, cry = 2.92 * PI;
  const double cu0 = 1.32, cus = 0.06, cux = 2.09 * PI, cuy = 3.12 * PI;
  const double cv0 = 1.18, cvs = 0.03, cvx = 2.15 * PI, cvy = 3.32 * PI;
  const double cp0 = 1.62, cps = 0.31, cpx = 3.79 * PI, cpy = 2.98 * PI;
  double g = o->op.gamma;
  double r = manufactured_sol(cr0, crs, crx, cry, 0, 0, xc, yc);
  double u = manufactured_sol(cu0, cus, cux, cuy, 0, 0, xc, yc);
  double v = manufactured_sol(cv0, cvs, cvx, cvy, 0, 0, xc, yc);
  double p = manufactured_sol(cp0, cps, cpx, cpy, 0, 0, xc, yc);
  sol[0] = r; sol[1] = u; sol[2] = v; sol[3] = p;
  if (!rhs) return;
  double rx = manufactured_sol(cr0, crs, crx, cry, 1, 0, xc, yc);
  double ux = manufactured_sol(cu0, cus, cux, cuy, 1, 0, xc, yc);
  double vx = manufactured_sol(cv0, cvs, cvx, cvy, 1, 0, xc, yc);
  double px = manufactured_sol(cp0, cps, cpx, cpy, 1, 0, xc, yc);
  double ry = manufactured_sol(cr0, crs, crx, cry, 0, 1, xc, yc);
  double uy = manufactured_sol(cu0, cus, cux, cuy, 0, 1, xc, yc);
  double vy = manufactured_sol(cv0, cvs, cvx, cvy, 0, 1, xc, yc);
  double py = manufactured_sol(cp0, cps, cpx, cpy, 0, 1, xc, yc);
  double rH = g / (g - 1.0) * p + r * u * u / 2.0 + r * v * v / 2.0;
  double rHx = g / (g - 1.0) * px + rx * (u * u + v * v) / 2.0 + r * (u * ux + v * vx);
  double rHy = g / (g - 1.0) * py + ry * (u * u + v * v) / 2.0 + r * (u * uy + v * vy);
  (void)uy;
  /* continuity as written in the reference (src/mms.f90:169): the second
   * term is u*rx, not r*ux */
  rhs[0] = rx * u + u * rx + ry * v + r * vy;
  rhs[1] = rx * u * u + 2.0 * r * u * ux + ry * u * v + r * uy * v + r * u * vy + px;
  rhs[2] = rx * u * v + r * ux * v + r * u * vx + ry * v * v + 2.0 * r * v * vy + py;
  rhs[3] = u * rHx + ux * rH + v * rHy + vy * rH;
}

void orc_vortex_state(const orc_t *o, double t, double x, double y, double *pv) {
  double g = o->op.gamma;
  double K = o->op.vortex_K;
  double rho_inf = o->op.vortex_inf[0], u_inf = o->op.vortex_inf[1];
  double v_inf = o->op.vortex_inf[2], p_inf = o->op.vortex_inf[3];
  double T_inf = p_inf / rho_inf;
  double xc = o->op.vortex_x0 + u_inf * t;
  double yc = o->op.vortex_y0 + v_inf * t;
  double dx = x - xc, dy = y - yc;
  double r = sqrt(dx * dx + dy * dy);
  double c = K / (2.0 * PI);
  double u = u_inf - c * dy * exp(0.5 * (1.0 - r * r));
  double v = v_inf + c * dx * exp(0.5 * (1.0 - r * r));
  double temp = T_inf - (c * c) * (g - 1.0) / (2.0 * g) * exp(1.0 - r * r);
  double rho = pow(temp, 1.0 / (g - 1.0));
  double p = pow(rho, g);
  pv[0] = rho; pv[1] = u; pv[2] = v; pv[3] = p;
}

/* bc_flux ghost state: src/residual.f90:183-255 */
int orc_bc_state(const orc_t *o, int bc_type, double t, double x, double y, double nx, double ny,
                 const double *pfL, double *pfR) {
  for (int k = 0; k < 4; ++k) pfR[k] = 0.0;
  if (bc_type == ORC_BC_FREESTREAM) {
    for (int k = 0; k < 4; ++k) pfR[k] = o->pvar_inf[k];
  } else if (bc_type == ORC_BC_SLIP_WALL) {
    for (int k = 0; k < 4; ++k) pfR[k] = pfL[k];
    double un = pfL[1] * nx + pfL[2] * ny;
#if ORC_MUTANT == MUT_SLIP_NO_FACTOR2
    pfR[1] = pfL[1] - 1.0 * un * nx;
    pfR[2] = pfL[2] - 1.0 * un * ny;
#else
    pfR[1] = pfL[1] - 2.0 * un * nx;
    pfR[2] = pfL[2] - 2.0 * un * ny;
#endif
  } else if (bc_type == ORC_BC_DIRICHLET) {
    if (o->op.lvortex) orc_vortex_state(o, t, x, y, pfR);
    else orc_mms_state(o, x, y, pfR, NULL);
  } else {
    set_err("Boundary condition type %d not implemented (src/residual.f90:206-220)", bc_type);
    return -1;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* compute_residual: src/residual.f90:23-177                                */
/* ------------------------------------------------------------------------ */
/* flux * area and ws * area of interior face i (position in edge_intr) */
static void interior_face(const orc_t *o, int i, double *fa, double *wa) {
  double g = o->op.gamma, kap = o->kappa;
  const double *gx = o->grad, *gy = o->grad + 4 * o->ncells;
  int ie = o->edge_intr[i];
  double xf = o->ex[ie], yf = o->ey[ie], af = o->ea[ie], nxf = o->enx[ie], nyf = o->eny[ie];
  int icL = o->ec1[ie], icR = o->ec2[ie];
  double xcL = o->cx[icL], ycL = o->cy[icL];
  double xcR = o->cx[icR] + o->esx[ie], ycR = o->cy[icR] + o->esy[ie];
  double pfL[4], pfR[4], flux[4], wsm;
  for (int k = 0; k < 4; ++k) {
    double gC = o->pvar[4 * icR + k] - o->pvar[4 * icL + k];
    double gL = (xf - xcL) * gx[4 * icL + k] + (yf - ycL) * gy[4 * icL + k];
    double gR = (xf - xcR) * gx[4 * icR + k] + (yf - ycR) * gy[4 * icR + k];
    pfL[k] = o->pvar[4 * icL + k] + o->phi[icL] * (kap / 2.0 * gC + (1.0 - kap) * gL);
    pfR[k] = o->pvar[4 * icR + k] + o->phi[icR] * (-(kap / 2.0 * gC) + (1.0 - kap) * gR);
  }
  orc_roe(g, pfL, pfR, nxf, nyf, flux, &wsm);
  for (int k = 0; k < 4; ++k) fa[k] = flux[k] * af;
  *wa = wsm * af;
}

/* flux * area and ws * area of boundary edge i of boundary ib; returns its cell */
static int boundary_face(const orc_t *o, int ib, int i, double time, double *fa, double *wa) {
  const double *gx = o->grad, *gy = o->grad + 4 * o->ncells;
  int ie = o->bnd_edges[o->bnd_eptr[ib] + i];
  double xf = o->ex[ie], yf = o->ey[ie], af = o->ea[ie], nxf = o->enx[ie], nyf = o->eny[ie];
  int icL = o->bnd_cells[o->bnd_cptr[ib] + i];
  double xcL = o->cx[icL], ycL = o->cy[icL];
  double pfL[4], pfR[4], flux[4], wsm;
  for (int k = 0; k < 4; ++k)
    pfL[k] = o->pvar[4 * icL + k] +
             o->phi[icL] * ((xf - xcL) * gx[4 * icL + k] + (yf - ycL) * gy[4 * icL + k]);
  orc_bc_state(o, o->bnd_type[ib], time, xf, yf, nxf, nyf, pfL, pfR);
  orc_roe(o->op.gamma, pfL, pfR, nxf, nyf, flux, &wsm);
  for (int k = 0; k < 4; ++k) fa[k] = flux[k] * af;
  *wa = wsm * af;
  return icL;
}

/* compute_residual: src/residual.f90:23-177.  Serially the face fluxes are
 * scattered to their cells in face order; threaded, every face's products
 * go to an array and each cell adds its faces in that same order, so both
 * give the same bits. */
void orc_compute_residual(orc_t *o, double time) {
  int N = o->ncells;
  orc_cvar2pvar(o);
  orc_gradient(o);
  orc_limiter(o);
  if (o->nthr > 1 && o->cf_ptr) {
    int nfi = o->nedges_intr;
    ORC_PAR
    for (int i = 0; i < nfi; ++i) interior_face(o, i, o->fa + 4 * (size_t)i, o->wa + i);
    for (int ib = 0; ib < o->nb; ++ib) {
      int b0 = nfi + (o->bnd_eptr[ib] - o->bnd_eptr[0]);
      ORC_PAR
      for (int i = 0; i < o->bnd_nedges[ib]; ++i)
        (void)boundary_face(o, ib, i, time, o->fa + 4 * (size_t)(b0 + i), o->wa + b0 + i);
    }
    ORC_PAR
    for (int ic = 0; ic < N; ++ic) {
      double r[4] = {0.0, 0.0, 0.0, 0.0}, w = 0.0;
      for (int e = o->cf_ptr[ic]; e < o->cf_ptr[ic + 1]; ++e) {
        int sg = o->cf[e], f = (sg > 0 ? sg : -sg) - 1;
        for (int k = 0; k < 4; ++k) r[k] = sg > 0 ? r[k] + o->fa[4 * (size_t)f + k] : r[k] - o->fa[4 * (size_t)f + k];
        w = w + o->wa[f];
      }
      for (int k = 0; k < 4; ++k) o->resid[4 * ic + k] = -r[k] / o->vol[ic];
      o->ws[ic] = w;
    }
    return;
  }
  memset(o->resid, 0, 4 * N * 8);
  memset(o->ws, 0, N * 8);
  for (int i = 0; i < o->nedges_intr; ++i) {
    int ie = o->edge_intr[i];
    int icL = o->ec1[ie], icR = o->ec2[ie];
    double fa[4], wa;
    interior_face(o, i, fa, &wa);
    for (int k = 0; k < 4; ++k) {
      o->resid[4 * icL + k] = o->resid[4 * icL + k] + fa[k];
      o->resid[4 * icR + k] = o->resid[4 * icR + k] - fa[k];
    }
    o->ws[icL] = o->ws[icL] + wa;
    o->ws[icR] = o->ws[icR] + wa;
  }
  for (int ib = 0; ib < o->nb; ++ib) {
    for (int i = 0; i < o->bnd_nedges[ib]; ++i) {
      double fa[4], wa;
      int icL = boundary_face(o, ib, i, time, fa, &wa);
      for (int k = 0; k < 4; ++k) o->resid[4 * icL + k] = o->resid[4 * icL + k] + fa[k];
      o->ws[icL] = o->ws[icL] + wa;
    }
  }
  for (int ic = 0; ic < N; ++ic)
    for (int k = 0; k < 4; ++k) o->resid[4 * ic + k] = -o->resid[4 * ic + k] / o->vol[ic];
}

/* Threads for the per-cell work of later calls (1 = serial, the default);
 * builds each cell's face list in the serial scatter order. */
int orc_set_threads(orc_t *o, int n) {
  if (!o || n < 1) return -1;
  o->nthr = n;
  if (n == 1 || o->cf_ptr) return 0;
  int N = o->ncells, nfi = o->nedges_intr, nbe = 0;
  for (int ib = 0; ib < o->nb; ++ib) nbe += o->bnd_nedges[ib];
  int *cnt = xcalloc(N + 1, sizeof(int));
  for (int i = 0; i < nfi; ++i) {
    int ie = o->edge_intr[i];
    cnt[o->ec1[ie] + 1]++;
    cnt[o->ec2[ie] + 1]++;
  }
  for (int ib = 0; ib < o->nb; ++ib)
    for (int i = 0; i < o->bnd_nedges[ib]; ++i) cnt[o->bnd_cells[o->bnd_cptr[ib] + i] + 1]++;
  for (int c = 0; c < N; ++c) cnt[c + 1] += cnt[c];
  o->cf_ptr = cnt;
  o->cf = xcalloc((size_t)cnt[N] + 1, sizeof(int));
  int *fill = xcalloc(N, sizeof(int));
  for (int i = 0; i < nfi; ++i) {
    int ie = o->edge_intr[i], a = o->ec1[ie], b = o->ec2[ie];
    o->cf[cnt[a] + fill[a]++] = i + 1;
    o->cf[cnt[b] + fill[b]++] = -(i + 1);
  }
  int f = nfi;
  for (int ib = 0; ib < o->nb; ++ib)
    for (int i = 0; i < o->bnd_nedges[ib]; ++i, ++f) {
      int c = o->bnd_cells[o->bnd_cptr[ib] + i];
      o->cf[cnt[c] + fill[c]++] = f + 1;
    }
  free(fill);
  o->fa = xcalloc(4 * (size_t)(nfi + nbe) + 4, sizeof(double));
  o->wa = xcalloc((size_t)(nfi + nbe) + 1, sizeof(double));
  return 0;
}

/* compute_local_time: src/runge_kutta.f90:424-437 */
void orc_local_time(orc_t *o) {
  ORC_PAR
  for (int ic = 0; ic < o->ncells; ++ic) o->dt_local[ic] = o->op.cfl * o->vol[ic] / (0.5 * o->ws[ic]);
}

/* initialize_solution: src/initialize.f90:19-90 (free stream / vortex / MMS) */
void orc_initialize(orc_t *o, int mode, double t0) {
  for (int ic = 0; ic < o->ncells; ++ic) {
    double *pv = &o->pvar[4 * ic];
    if (mode == 0) {
      orc_mms_state(o, o->cx[ic], o->cy[ic], pv, NULL);
    } else {
      for (int k = 0; k < 4; ++k) pv[k] = o->pvar_inf[k];
      if (o->op.lvortex) orc_vortex_state(o, t0, o->cx[ic], o->cy[ic], pv);
    }
  }
  orc_pvar2cvar(o);
}

/* error_isentropic_vortex: src/mms.f90:271-365 */
void orc_vortex_error(const orc_t *o, double time, double *out) {
  double g = o->op.gamma;
  double mx[4] = {0, 0, 0, 0}, l1[4] = {0, 0, 0, 0}, l2[4] = {0, 0, 0, 0}, eL2 = 0.0;
  /* erho (src/mms.f90:292,321) is zero outside the interior cells; its
   * maxloc (363) is the first cell holding the largest value */
  int imax = 0;
  double emax = 0.0;
  for (int i = 0; i < o->ncells_intr; ++i) {
    int ic = o->cell_intr[i];
    double pv[4];
    orc_vortex_state(o, time, o->cx[ic], o->cy[ic], pv);
    double ex[4];
    ex[0] = pv[0];
    ex[1] = pv[0] * pv[1];
    ex[2] = pv[0] * pv[2];
    ex[3] = pv[3] / (g - 1.0) + 0.5 * ex[0] * (pv[1] * pv[1] + pv[2] * pv[2]);
    double d[4];
    for (int k = 0; k < 4; ++k) d[k] = fabs(o->cvar[4 * ic + k] - ex[k]);
    if (d[0] > emax || (d[0] == emax && emax > 0.0 && ic < imax)) { emax = d[0]; imax = ic; }
    for (int k = 0; k < 4; ++k) {
      mx[k] = (d[k] > mx[k]) ? d[k] : mx[k];
      l1[k] = l1[k] + d[k];
      l2[k] = l2[k] + d[k] * d[k];
    }
    eL2 = eL2 + d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + d[3] * d[3];
  }
  double n = (double)o->ncells_intr;
  for (int k = 0; k < 4; ++k) {
    out[3 * k] = mx[k];
    out[3 * k + 1] = l1[k] / n;
    out[3 * k + 2] = sqrt(l2[k] / n);
  }
  out[12] = sqrt(eL2 / n);
  out[13] = o->cx[imax];
  out[14] = o->cy[imax];
}

/* ------------------------------------------------------------------------ */
/* time integration: src/runge_kutta.f90:25-418 (+ builder-new SSP-RK3)     */
/* ------------------------------------------------------------------------ */
static void residual_norms(const orc_t *o, const double *cvar0, double *out4) {
  int N = o->ncells;
  for (int k = 0; k < 4; ++k) {
    double s = 0.0;
    for (int ic = 0; ic < N; ++ic) {
      double d = fabs(o->cvar[4 * ic + k] - cvar0[4 * ic + k]);
      s = s + d * d;
    }
    out4[k] = sqrt(s / (double)N);
  }
}

void orc_time_integration(orc_t *o, double t1, int nsteps, double *norms, double *verr) {
  int N = o->ncells;
  double dt = o->op.dt;
  double *cvar0 = xcalloc(4 * N, 8), *fcvar = xcalloc(4 * N, 8);
  /* runge_kutta_init 36-65 */
  double ak = 1, bk = 2, ck = 2, dk = 1;
  switch (o->op.rk_order) {
    case 1: ak = 3.60897; bk = 2.04; ck = 0.34206; dk = 0.00897; break;
    case 2: ak = 0.11; bk = 3.92; ck = 1.86; dk = 0.11; break;
    case 3: ak = 0.65; bk = 2.7; ck = 2.0; dk = 0.65; break;
    default: break;
  }
#if ORC_MUTANT == MUT_RK2_SWAP_BC
  if (o->op.rk_order == 2) { double tmp = bk; bk = ck; ck = tmp; }
#elif ORC_MUTANT == MUT_RK1_SWAP_AD
  if (o->op.rk_order == 1) { double tmp = ak; ak = dk; dk = tmp; }
#endif
  double h_rk[4] = {dt / 2.0, dt / 2.0, dt, dt / 6.0}, rk_coef[4] = {ak, bk, ck, dk};
#if ORC_MUTANT == MUT_RK3_H3
  if (o->op.rk_order == 3) h_rk[2] = dt / 2.0;
#endif
  double dts[4] = {0, 0, 0, 0}, dte[4] = {0, 0, 0, 0};
  if (o->op.rk_scheme == ORC_RK_SSP42) {
    h_rk[0] = h_rk[1] = h_rk[2] = dt / 3.0; h_rk[3] = dt / 4.0;
    rk_coef[0] = rk_coef[1] = rk_coef[2] = rk_coef[3] = 1.0;
    dts[0] = 0.0;            dte[0] = dt / 3.0;
    dts[1] = dt / 3.0;       dte[1] = dt * 2.0 / 3.0;
    dts[2] = dt * 2.0 / 3.0; dte[2] = dt;
    dts[3] = dt;             dte[3] = dt;
  }
  for (int istep = 1; istep <= nsteps; ++istep) {
    double t = t1 + (double)(istep - 1) * dt;
    double told = t, tend = t + dt;
    memcpy(cvar0, o->cvar, 4 * N * 8);
    memset(fcvar, 0, 4 * N * 8);
    if (o->op.rk_scheme == ORC_RK_CLASSIC) {
      double t12 = told + 0.5 * dt, tnew = told + dt;
      double tstart[4] = {told, t12, t12, tnew};
      tend = tnew;
      for (int rk = 0; rk < 4; ++rk) {
        orc_compute_residual(o, tstart[rk]);
        if (o->op.steady && rk == 0) orc_local_time(o);
        ORC_PAR
        for (int i = 0; i < 4 * N; ++i) fcvar[i] = fcvar[i] + rk_coef[rk] * o->resid[i];
        if (!o->op.steady) {
          /* time_integ_RK 156-162 */
          if (rk < 3) {
            ORC_PAR
            for (int i = 0; i < 4 * N; ++i) o->cvar[i] = cvar0[i] + h_rk[rk] * o->resid[i];
          } else {
            ORC_PAR
            for (int i = 0; i < 4 * N; ++i) o->cvar[i] = cvar0[i] + h_rk[rk] * fcvar[i];
          }
        } else {
          /* time_integ_RK_steady 299-313 (the reference aborts before its
           * first step, src/runge_kutta.f90:270,321; restated as written) */
          double cst = 1.0 / 2.0;
          if (rk == 2) cst = 1.0;
          if (rk == 3) cst = 1.0 / 6.0;
          ORC_PAR
          for (int ic = 0; ic < N; ++ic)
            for (int k = 0; k < 4; ++k) {
              int i = 4 * ic + k;
              o->cvar[i] = cvar0[i] + o->dt_local[ic] * cst * (rk < 3 ? o->resid[i] : fcvar[i]);
            }
        }
      }
    } else if (o->op.rk_scheme == ORC_RK_SSP42) {
      for (int rk = 0; rk < 4; ++rk) {
        double ts = told + dts[rk];
        tend = told + dte[rk];
        orc_compute_residual(o, ts);
        if (!o->op.steady) {
          /* time_integ_SSPRK 225-226 */
          ORC_PAR
          for (int i = 0; i < 4 * N; ++i) o->cvar[i] = cvar0[i] + h_rk[rk] * (rk_coef[rk] * o->resid[i] + fcvar[i]);
        } else {
          /* time_integ_SSPRK_steady 379-385 */
          if (rk == 0) orc_local_time(o);
          double cst = (rk == 3) ? 1.0 / 4.0 : 1.0 / 3.0;
          ORC_PAR
          for (int ic = 0; ic < N; ++ic)
            for (int k = 0; k < 4; ++k) {
              int i = 4 * ic + k;
              o->cvar[i] = cvar0[i] + o->dt_local[ic] * cst * (rk_coef[rk] * o->resid[i] + fcvar[i]);
            }
        }
        ORC_PAR
        for (int i = 0; i < 4 * N; ++i) fcvar[i] = fcvar[i] + o->resid[i];
      }
    } else {
      /* builder-new SSP-RK3 (Shu-Osher): u_k = a_k*u0 + b_k*(u + dt*R(u)),
       * (a,b) = (0,1), (3/4,1/4), (1/3,2/3); stage times t, t+dt, t+dt/2 */
      static const double A[3] = {0.0, 0.75, 1.0 / 3.0}, B[3] = {1.0, 0.25, 2.0 / 3.0};
      double ts[3] = {told, told + dt, told + 0.5 * dt};
      for (int s = 0; s < 3; ++s) {
        orc_compute_residual(o, ts[s]);
        if (o->op.steady && s == 0) orc_local_time(o);
        ORC_PAR
        for (int ic = 0; ic < N; ++ic) {
          double h = o->op.steady ? o->dt_local[ic] : dt;
          for (int k = 0; k < 4; ++k) {
            int i = 4 * ic + k;
            o->cvar[i] = A[s] * cvar0[i] + B[s] * (o->cvar[i] + h * o->resid[i]);
          }
        }
      }
      tend = told + dt;
    }
    if (verr) orc_vortex_error(o, tend, verr + 15 * (istep - 1));
    if (norms) residual_norms(o, cvar0, norms + 4 * (istep - 1));
  }
  free(cvar0); free(fcvar);
}

/* write_inst_cp_un C_L / C_D part: src/io.f90:340-449 (summed over all
 * slip_wall / solid_wall boundaries; one wall in every shipped case) */
void orc_wall_forces(const orc_t *o, double *out) {
  int N = o->ncells;
  double *pc = xcalloc(N, 8), *dp = xcalloc(2 * N, 8);
  for (int ic = 0; ic < N; ++ic) pc[ic] = o->pvar[4 * ic + 3];
  orc_gradient_1var(o, pc, dp);
  double M = o->op.mach, pinf = o->pvar_inf[3];
  double cn = 0, ca = 0, cn1 = 0, ca1 = 0;
  for (int ib = 0; ib < o->nb; ++ib) {
    if (o->bnd_type[ib] != ORC_BC_SLIP_WALL && o->bnd_type[ib] != ORC_BC_SOLID_WALL) continue;
    for (int i = 0; i < o->bnd_nedges[ib]; ++i) {
      int ie = o->bnd_edges[o->bnd_eptr[ib] + i];
      int ic = o->bnd_cells[o->bnd_cptr[ib] + i];
      double af = o->ea[ie], nx = o->enx[ie], ny = o->eny[ie];
      double pw = o->pvar[4 * ic + 3] + (o->ex[ie] - o->cx[ic]) * dp[ic] + (o->ey[ie] - o->cy[ic]) * dp[N + ic];
      double cp = 2.0 / (M * M) * (pw - pinf);
      double cp1 = 2.0 / (M * M) * (o->pvar[4 * ic + 3] - pinf);
      cn = cn + cp * ny * af;
      ca = ca + cp * nx * af;
      cn1 = cn1 + cp1 * ny * af;
      ca1 = ca1 + cp1 * nx * af;
    }
  }
  double al = o->op.aoa_deg;
  out[0] = cn * cosd(al) - ca * sind(al);
  out[1] = cn * sind(al) + ca * cosd(al);
  out[2] = cn1 * cosd(al) - ca1 * sind(al);
  out[3] = cn1 * sind(al) + ca1 * cosd(al);
  free(pc); free(dp);
}

/* Wall pressure coefficient and normal velocity per wall face,
 * src/io.f90:380-418: out[4k..4k+3] = x_f, cp (face-extrapolated pressure),
 * cp (cell pressure), V_n = (u_f, v_f).n with face-extrapolated velocity;
 * un3 = max |V_n|, sqrt(mean V_n^2), mean |V_n| over the wall faces in
 * boundary order (the reference's per-boundary figures when there is one
 * wall boundary).  Returns the number of wall faces. */
int orc_wall_cp(const orc_t *o, double *out, double *un3) {
  int N = o->ncells;
  double *q = xcalloc(N, 8), *dp = xcalloc(2 * N, 8), *du = xcalloc(2 * N, 8), *dv = xcalloc(2 * N, 8);
  for (int ic = 0; ic < N; ++ic) q[ic] = o->pvar[4 * ic + 3];
  orc_gradient_1var(o, q, dp);
  for (int ic = 0; ic < N; ++ic) q[ic] = o->pvar[4 * ic + 1];
  orc_gradient_1var(o, q, du);
  for (int ic = 0; ic < N; ++ic) q[ic] = o->pvar[4 * ic + 2];
  orc_gradient_1var(o, q, dv);
  double M = o->op.mach, pinf = o->pvar_inf[3];
  double un_max = 0.0, un_l2 = 0.0, un_l1 = 0.0;
  int k = 0;
  for (int ib = 0; ib < o->nb; ++ib) {
    if (o->bnd_type[ib] != ORC_BC_SLIP_WALL && o->bnd_type[ib] != ORC_BC_SOLID_WALL) continue;
    for (int i = 0; i < o->bnd_nedges[ib]; ++i) {
      int ie = o->bnd_edges[o->bnd_eptr[ib] + i];
      int ic = o->bnd_cells[o->bnd_cptr[ib] + i];
      double ddx = o->ex[ie] - o->cx[ic], ddy = o->ey[ie] - o->cy[ic];
      double pw = o->pvar[4 * ic + 3] + ddx * dp[ic] + ddy * dp[N + ic];
      double uw = o->pvar[4 * ic + 1] + ddx * du[ic] + ddy * du[N + ic];
      double vw = o->pvar[4 * ic + 2] + ddx * dv[ic] + ddy * dv[N + ic];
      double un = uw * o->enx[ie] + vw * o->eny[ie];
      out[4 * k] = o->ex[ie];
      out[4 * k + 1] = 2.0 / (M * M) * (pw - pinf);
      out[4 * k + 2] = 2.0 / (M * M) * (o->pvar[4 * ic + 3] - pinf);
      out[4 * k + 3] = un;
      un_l2 = un_l2 + un * un;
      un_l1 = un_l1 + fabs(un);
      if (fabs(un) > un_max) un_max = fabs(un);
      ++k;
    }
  }
  un3[0] = un_max;
  un3[1] = k ? sqrt(un_l2 / (double)k) : 0.0;
  un3[2] = k ? un_l1 / (double)k : 0.0;
  free(q); free(dp); free(du); free(dv);
  return k;
}

/* mms_init source terms at cell centres: src/mms.f90:106-115 */
void orc_mms_source(const orc_t *o, double *src) {
  for (int ic = 0; ic < o->ncells; ++ic) {
    double sol[4];
    orc_mms_state(o, o->cx[ic], o->cy[ic], sol, &src[4 * ic]);
  }
}
