/*
 * fvs_oracle.h -- CPU restatement of the fvs2d residual / time-step path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only for the results
 * the tests compare with (or as the timed CPU baseline).  The product (fvs2d_amd, libfvs2d_hip.so) never
 * links, loads or calls it.
 *
 * Every routine restates one reference routine, cited file:line against the
 * Fortran sources of shirzadgit/fvs2d.  Indices are 0-based; "no cell" is -1
 * (the reference uses 1-based indices and 0 for "no cell").
 */
#ifndef FVS_ORACLE_H
#define FVS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* boundary types, src/residual.f90:194-221 (+ builder-new periodic handled
 * through node identification, not as a boundary) */
enum { ORC_BC_FREESTREAM = 0, ORC_BC_SLIP_WALL = 1, ORC_BC_DIRICHLET = 2, ORC_BC_SOLID_WALL = 3 };

/* time integrators: src/runge_kutta.f90:94-114 (+ builder-new SSP-RK3) */
enum { ORC_RK_CLASSIC = 0, ORC_RK_SSP42 = 1, ORC_RK_SSP3 = 2 };

typedef struct {
  int grad_method;     /* 1 GGCB, 2 GGNB, 3 LSQ          (src/input.f90:204-228) */
  int lsq_nn;          /* 1 node-neighbour stencil, 0 face-neighbour        */
  double lsq_p;        /* LSQ weight power p, w = 1/d^p                      */
  int limiter;         /* 0 none, 1 Venkatakrishnan, 2 Barth, 3 Van Albada  */
  int recon;           /* 1 upwind-1st, 2 upwind-2nd, 3 UMUSCL              */
  double kappa;        /* umuscl_cst; forced to 0 for recon 1 and 2         */
  double gamma, mach, aoa_deg;
  int lvortex;         /* Dirichlet boundaries use the isentropic vortex    */
  double vortex_x0, vortex_y0, vortex_K;
  double vortex_inf[4];/* rho, u, v, p of the vortex free stream            */
  int rk_scheme;       /* ORC_RK_*                                           */
  int rk_order;        /* 1..4 for ORC_RK_CLASSIC                           */
  int steady;          /* local time stepping                               */
  double dt, cfl;
} orc_opts;

typedef struct orc orc_t;

/* Build mesh topology, geometry, stencils and solution arrays.
 * cell_nodes: ntri*3 then nquad*4 node ids (0-based), CCW.
 * node_canon: NULL, or per node the canonical id used for topology (periodic
 *             identification; builder-new).
 * bnd_cells: concatenated 0-based cell lists of the nb boundaries.
 * Returns NULL on error; orc_last_error() then holds the message. */
orc_t *orc_create(int nnodes, const double *xy, int ntri, int nquad, const int *cell_nodes,
                  const int *node_canon, int nb, const int *bnd_ncells, const int *bnd_types,
                  const int *bnd_cells, const orc_opts *opts);
void orc_destroy(orc_t *o);
const char *orc_last_error(void);

/* sizes: [ncells, nedges, nnodes, nedges_intr, nedges_bndr, ncells_intr, ncells_bndr, nbndries] */
void orc_sizes(const orc_t *o, int *out8);
/* edges: n1,n2,c1,c2 (int, nedges each) and x,y,area,nx,ny,sx,sy (double, nedges each) */
void orc_get_edges(const orc_t *o, int *n1, int *n2, int *c1, int *c2, double *geom7);
/* cells: x,y,vol (3*ncells), nvrt, nghbre[4*ncells], edge[4*ncells], nrmlsign[4*ncells] */
void orc_get_cells(const orc_t *o, double *xyv, int *nvrt, int *nghbre, int *edge, double *nrmlsign);
void orc_get_lists(const orc_t *o, int *edge_intr, int *cell_intr, int *bnd_nedges, int *bnd_edges);
/* LSQ stencil in CSR form: ptr (ncells+1), cell, w, coef (2 per entry) */
int  orc_lsq_nnz(const orc_t *o);
void orc_get_lsq(const orc_t *o, int *ptr, int *cell, double *w, double *coef);

/* state access (cvar, pvar: 4*ncells AoS like the Fortran (4,N) arrays) */
void orc_set_cvar(orc_t *o, const double *cvar);
void orc_get_cvar(const orc_t *o, double *cvar);
void orc_get_pvar(const orc_t *o, double *pvar);
void orc_set_pvar(orc_t *o, const double *pvar);
/* grad: [2][ncells][4] like the Fortran grad(4,N,2) */
void orc_get_grad(const orc_t *o, double *grad);
void orc_get_phi(const orc_t *o, double *phi);
void orc_get_resid(const orc_t *o, double *resid);
void orc_get_ws(const orc_t *o, double *ws);
void orc_get_dt_local(const orc_t *o, double *dtl);

/* operators */
void orc_cvar2pvar(orc_t *o);
void orc_pvar2cvar(orc_t *o);
void orc_gradient(orc_t *o);
void orc_gradient_1var(const orc_t *o, const double *fc, double *dfc /* [2][ncells] */);
void orc_limiter(orc_t *o);
void orc_compute_residual(orc_t *o, double time);
void orc_local_time(orc_t *o);
void orc_roe(double gamma, const double *pL, const double *pR, double nx, double ny, double *flux, double *ws_max);
double orc_limiter_fn(int limiter_type, double a, double b, double vol);
void orc_vortex_state(const orc_t *o, double t, double x, double y, double *pv);
void orc_mms_state(const orc_t *o, double x, double y, double *sol, double *rhs);
int  orc_bc_state(const orc_t *o, int bc_type, double t, double x, double y, double nx, double ny,
                  const double *pfL, double *pfR);

/* initial solution: mode 1 free stream (+vortex if lvortex), 0 MMS; t0 = start time */
void orc_initialize(orc_t *o, int mode, double t0);
/* integrate nsteps from t1; norms: nsteps*4 (may be NULL); verr: nsteps*15 (may be NULL) */
void orc_time_integration(orc_t *o, double t1, int nsteps, double *norms, double *verr);
/* per-cell work of later calls on n threads (OpenMP; 1 = serial, the
 * default); the results are bit-identical to the serial ones */
int  orc_set_threads(orc_t *o, int n);
/* the 13 norms of src/mms.f90:351-361, then x, y of the cell with the
 * largest density error (src/mms.f90:363) */
void orc_vortex_error(const orc_t *o, double t, double *out15);
/* wall forces over slip/solid walls: cl, cd, cl1, cd1 (src/io.f90:340-449) */
void orc_wall_forces(const orc_t *o, double *out4);
int orc_wall_cp(const orc_t *o, double *out, double *un3);
void orc_mms_source(const orc_t *o, double *src /* 4*ncells */);

#ifdef __cplusplus
}
#endif
#endif
