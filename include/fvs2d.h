/*
 * fvs2d.h -- C-ABI of the MI355X-native fvs2d residual / time-step path.
 *
 * The reference (shirzadgit/fvs2d) exposes this path as Fortran module
 * procedures working on module-global arrays; a Fortran driver reaches this
 * library through the iso_c_binding modules in fvs2d_amd/fortran/, which keep
 * the reference's module and procedure names.  Each entry point below names
 * the reference procedure it replaces.
 *
 * Conventions
 *   - plain pointers and sizes only; all reals are double, all ints int32;
 *   - cell, node and edge ids are 0-based (the Fortran side subtracts 1);
 *   - per-cell arrays are in the caller's (reference) cell order; the library
 *     may store cells in another order on the device and permutes at the
 *     boundary;
 *   - cvar / pvar / resid are (4, ncells) column-major, i.e. C double[N][4]
 *     like the Fortran cvar(4,N); grad is double[2][N][4] like grad(4,N,2);
 *   - every function returns 0 on success and a negative code on error; the
 *     message is then available from fvs_last_error(); no C++ exception
 *     leaves the library (host memory exhausted during preparation or a
 *     transfer returns FVS_ENOMEM, any other host failure FVS_ESTATE);
 *   - one context owns one HIP stream and is not re-entrant; calls on one
 *     context must come from one host thread at a time.
 */
#ifndef FVS2D_H
#define FVS2D_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fvs_ctx fvs_ctx;

/* boundary types: src/residual.f90:194-221 */
enum { FVS_BC_FREESTREAM = 0, FVS_BC_SLIP_WALL = 1, FVS_BC_DIRICHLET = 2, FVS_BC_SOLID_WALL = 3 };

/* time integrators: src/runge_kutta.f90:94-114; FVS_RK_SSP3 has no reference
 * counterpart (Shu-Osher SSP-RK3).  FVS_RK_SSP42_SO is the reference's
 * SSP-RK(4,2) (rk_coef 1, stage steps h/3, h/3, h/3, h/4,
 * src/runge_kutta.f90:56-60) evaluated in its two-register Shu-Osher form
 * u_k = u_{k-1} + h/3 R(u_{k-1}) (k = 1..3), u_4 = u_0/4 + 3/4 (u_3 + h/3
 * R(u_3)) instead of the reference's u_k = u_0 + h c_k (R + sum of earlier
 * R): the same method, equal in exact arithmetic, with its own rounding and
 * no accumulator, so its stages can run on conservative states (no
 * reference counterpart for the form) */
enum { FVS_RK_CLASSIC = 0, FVS_RK_SSP42 = 1, FVS_RK_SSP3 = 2, FVS_RK_SSP42_SO = 3 };

/* error codes */
enum { FVS_OK = 0, FVS_EINVAL = -1, FVS_EHIP = -2, FVS_ESTATE = -3, FVS_ENOMEM = -4 };

/* Scheme flags with the meaning src/input.f90:181-277 gives them. */
typedef struct {
  int grad_method;      /* 1 GGCB, 2 GGNB, 3 LSQ                                */
  int lsq_nn;           /* LSQ stencil: 1 node neighbours ("nn"), 0 face ("fn") */
  double lsq_p;         /* LSQ weight power: w = 1/d^p                          */
  int limiter;          /* 0 none, 1 Venkatakrishnan, 2 Barth-Jespersen, 3 Van Albada */
  int recon;            /* 1 upwind-1st, 2 upwind-2nd, 3 UMUSCL                 */
  double kappa;         /* UMUSCL constant (ignored unless recon == 3)          */
  double gamma, mach, aoa_deg;
  int lvortex;          /* Dirichlet boundaries carry the isentropic vortex (else MMS) */
  double vortex_x0, vortex_y0, vortex_K, vortex_inf[4]; /* fvs2d.vortex contents */
  int rk_scheme;        /* FVS_RK_*                                             */
  int rk_order;         /* 1..4, FVS_RK_CLASSIC only                            */
  int steady;           /* local time stepping                                  */
  double dt, cfl;
  int reorder;          /* device cell order: 0 reference order, 1 Morton (Z-curve)
                           blocks, 2 recursive coordinate bisection (compact
                           blocks on graded meshes); results are the same  */
  int kernel_path;      /* 0: fastest available (fused LDS flux+update, centroid LSQ;
                              SSP-RK3 stages read conservative stage states and
                              write no primitive state, with the LSQ gradient or an
                              unlimited Green-Gauss one);
                           1: face-flux buffer kernels only;
                           2: whole stage in one kernel (gradients recomputed per
                              block from LDS; LSQ, centroid form, triangle-sized
                              blocks; falls back to 0 where it does not apply);
                           3: as 0 without the per-block halo data (the LSQ
                              gradient gathers its stencil cell by cell) and with
                              SSP-RK3 stages on primitive states */
} fvs_options;

const char *fvs_last_error(void);

/* Mesh + scheme set-up: replaces grid_procs_init's topology part
 * (src/grid_procs.f90:31-59, 170-794), data_solution_init
 * (src/data_solution.f90:41-66), gradient_init (src/gradient.f90:18-38) and
 * runge_kutta_init (src/runge_kutta.f90:25-88).
 *   xy          2*nnodes node coordinates
 *   cell_nodes  ntri*3 then nquad*4 node ids, counter-clockwise
 *   node_canon  NULL, or per node the id it is identified with (periodic
 *               seams; no reference counterpart)
 *   bnd_*       the .bc file: per boundary its cell count and type, then the
 *               concatenated cell lists */
int fvs_create(fvs_ctx **out, int nnodes, const double *xy, int ntri, int nquad, const int *cell_nodes,
               const int *node_canon, int nb, const int *bnd_ncells, const int *bnd_types,
               const int *bnd_cells, const fvs_options *opt);
/* Batched instances (no reference counterpart; the reference solves one
 * case per run): ninst independent meshes concatenated into one (instance k
 * owns cells [cell_offsets[k], cell_offsets[k+1]) of the concatenation, its
 * boundaries listed with the others), each with its own free stream
 * (mach_aoa[2k], mach_aoa[2k+1] in degrees).  cell_offsets must increase
 * strictly from 0 to ntri + nquad (every instance has cells).  Every RK
 * stage is one launch over all instances; per-instance residual norms, masks
 * and forces below. */
int fvs_create_batch(fvs_ctx **out, int ninst, const int *cell_offsets, const double *mach_aoa, int nnodes,
                     const double *xy, int ntri, int nquad, const int *cell_nodes, int nb, const int *bnd_ncells,
                     const int *bnd_types, const int *bnd_cells, const fvs_options *opt);
/* nsteps steps of all active instances; inst_norms (nsteps*ninst*4, may be
 * NULL) receives each instance's log_res.plt norms per step */
int fvs_batch_time_integration(fvs_ctx *ctx, double t1, int nsteps, double *inst_norms);
/* active[k] != 0 keeps instance k stepping; the others are left out of the
 * launches (blocks holding only inactive instances are skipped, and a block
 * shared with an active instance leaves their cells untouched) */
int fvs_batch_set_active(fvs_ctx *ctx, const int *active);
/* C_L, C_D, C_L(cell), C_D(cell) per instance (4*ninst), src/io.f90:340-449 */
int fvs_batch_wall_forces(fvs_ctx *ctx, double *out);
/* Launch mode, no reference counterpart: on != 0 captures the launches of one
 * time step (two for three-stage schemes) in a HIP graph and replays it; the
 * instantiated graph is kept for later fvs_time_integration calls until a call
 * changes what a step launches (time step, limiter freezing, active
 * instances, autotuned choices).  Time-dependent boundary data (the
 * Dirichlet vortex) always use stream launches. */
int fvs_set_graph(fvs_ctx *ctx, int on);
/* Phase timers, the counterpart of the reference's cput_grad, cput_lim,
 * cput_flux and cput_rk (src/mainparam.f90:15-16, accumulated around each
 * phase and printed by src/fvs2d.f90:175-182): while on != 0, every
 * fvs_time_integration call records HIP events around each launch (and
 * replays no graph) and adds each launch's device time to its phase. */
int fvs_set_phase_timers(fvs_ctx *ctx, int on);
/* Seconds accumulated by the phase timers: sec4 = gradient (the limiter is
 * evaluated inside the gradient launches, so its time is here), limiter
 * (always 0 for that reason), face fluxes (with the stage update where it is
 * fused into the flux launch), Runge-Kutta (separate stage-update launches
 * and the residual-norm reduction). */
int fvs_phase_times(const fvs_ctx *ctx, double *sec4);
/* Launch mode, no reference counterpart: on != 0 runs each
 * fvs_time_integration call (or batch call) as one persistent cooperative
 * launch -- workgroups resident on every compute unit walk the launch blocks
 * (the active blocks of a masked batch) through each whole RK stage, meet at a
 * grid barrier after every stage and reduce each step's residual norms in the
 * same launch -- instead of per-stage launches.  Results are bit-identical to
 * them.  Needs the whole-stage kernel (kernel_path 2); time-dependent boundary
 * data and calls that want vortex errors use per-stage launches.  A barrier
 * wait is bounded: a launch whose workgroups fail to meet returns FVS_EHIP. */
int fvs_set_persistent(fvs_ctx *ctx, int on);
/* replaces data_sol_close / data_grid_close (src/data_solution.f90:130-140) */
int fvs_destroy(fvs_ctx *ctx);

/* [ncells, nedges, nnodes, nedges_intr, nedges_bndr, ncells_intr, ncells_bndr, nbndries] */
int fvs_sizes(const fvs_ctx *ctx, int *out8);

/* Host-only mesh preparation (no device needed): the topology, geometry and
 * stencils fvs_create builds, for callers that want the reference's grid data
 * (grid_procs, src/grid_procs.f90:170-794; stencil set-ups
 * src/gradient_lsq.f90:70-365) without a solver context. */
typedef struct fvs_mesh fvs_mesh;
int fvs_mesh_create(fvs_mesh **out, int nnodes, const double *xy, int ntri, int nquad, const int *cell_nodes,
                    const int *node_canon, int nb, const int *bnd_ncells, const int *bnd_types,
                    const int *bnd_cells, int grad_method, int lsq_nn, double lsq_p, int limiter);
/* The same for a batch of ninst instances concatenated as fvs_create_batch
 * takes them (cells of one kind; each instance owning a consecutive node range
 * and consecutive boundaries): every instance is prepared on its own, on
 * nthreads host threads (0: the cores available, at most 16), and the results
 * merged -- the numbering the preparation of the whole concatenation gives,
 * with stencil searches kept inside each instance.  fvs_create_batch prepares
 * its meshes this way.  No reference counterpart (the reference prepares one
 * mesh per run, src/grid_procs.f90:170-794). */
int fvs_mesh_create_batch(fvs_mesh **out, int ninst, const int *cell_offsets, int nnodes, const double *xy,
                          int ntri, int nquad, const int *cell_nodes, int nb, const int *bnd_ncells,
                          const int *bnd_types, const int *bnd_cells, int grad_method, int lsq_nn, double lsq_p,
                          int limiter, int nthreads);
int fvs_mesh_destroy(fvs_mesh *mesh);
/* the mesh a solver context was built on (owned by the context) */
const fvs_mesh *fvs_ctx_mesh(const fvs_ctx *ctx);
int fvs_mesh_sizes(const fvs_mesh *mesh, int *out8);
/* Topology as grid_data builds it (src/grid_procs.f90:401-791), reference
 * numbering: edges n1,n2,c1,c2 (c2 = -1 on boundaries) and geometry
 * x,y,area,nx,ny,sx,sy (7*nedges; sx,sy = periodic shift of c2, else 0);
 * cells x,y,vol (3*ncells), nghbre and edge (4*ncells, -1 padded).  Any
 * output may be NULL. */
int fvs_mesh_topology(const fvs_mesh *mesh, int *n1, int *n2, int *c1, int *c2, double *edge_geom7,
                      double *cell_xyv3, int *nghbre4, int *cell_edge4);
/* interior edge list (ascending), interior cell list, per-boundary edge
 * counts and the concatenated boundary edge lists (src/grid_procs.f90:697-791) */
int fvs_mesh_lists(const fvs_mesh *mesh, int *edge_intr, int *cell_intr, int *bnd_nedges, int *bnd_edges);
/* LSQ stencil (src/gradient_lsq.f90:16-365) in CSR form: ptr[ncells+1]; the
 * caller sizes cell/w/coef from fvs_mesh_lsq_nnz; coef holds 2 per entry. */
int fvs_mesh_lsq_nnz(const fvs_mesh *mesh);
int fvs_mesh_lsq(const fvs_mesh *mesh, int *ptr, int *cell, double *w, double *coef);
/* The device cell order fvs_create uses for a reorder option (0 reference,
 * 1 Morton, 2 recursive coordinate bisection): perm[i] is the reference id of
 * the cell at device position i (ncells entries).  For a batch mesh pass its
 * ninst and cell_offsets (instances stay contiguous and in order), else 0 and
 * NULL.  No reference counterpart: the reference keeps the input order
 * (src/grid_procs.f90:170-794); this shows callers the order the device
 * blocks are cut from. */
int fvs_mesh_cell_order(const fvs_mesh *mesh, int reorder, int ninst, const int *cell_offsets, int *perm);

/* cvar in / out (module array data_solution::cvar) */
int fvs_set_state(fvs_ctx *ctx, const double *cvar);
int fvs_get_state(fvs_ctx *ctx, double *cvar);
/* initialize_solution (src/initialize.f90:19-90): mode 1 free stream
 * (+ vortex at t0 if lvortex), mode 0 manufactured solution */
int fvs_initialize(fvs_ctx *ctx, int mode, double t0);

/* cvar2pvar (src/data_solution.f90:72-86); pvar may be NULL */
int fvs_primitives(fvs_ctx *ctx, double *pvar);
/* compute_gradient_cellcntr (src/gradient.f90:44-69) on the current pvar;
 * grad may be NULL */
int fvs_gradient(fvs_ctx *ctx, double *grad);
/* gradient_cellcntr_1var (src/gradient.f90:75-96): dfc is double[2][N] */
int fvs_gradient_1var(fvs_ctx *ctx, const double *fc, double *dfc);
/* compute_gradient_limiter (src/gradient_limiter.f90:19-97) on the current
 * pvar and grad; phi may be NULL */
int fvs_limiter(fvs_ctx *ctx, double *phi);
/* compute_residual (src/residual.f90:23-177): resid (dQ/dt) and ws_nrml;
 * either output may be NULL */
int fvs_compute_residual(fvs_ctx *ctx, double time, double *resid, double *ws_nrml);
/* compute_flux_invscid (src/flux_invscid.f90:17-32) on n face states:
 * pL,pR 4*n primitive states, nxy 2*n unit normals; flux 4*n, ws n */
int fvs_roe_flux_batch(int n, double gamma, const double *pL, const double *pR, const double *nxy,
                       double *flux, double *ws);
/* compute_local_time (src/runge_kutta.f90:424-437) from the last residual's
 * ws_nrml; dt may be NULL */
int fvs_local_dt(fvs_ctx *ctx, double *dt);
/* time_integration (src/runge_kutta.f90:94-114): nsteps steps from t1.
 * norms (nsteps*4, may be NULL) receives the log_res.plt norms
 * (src/runge_kutta.f90:169-184); verr (nsteps*15, may be NULL) after each
 * step the error_isentropic_vortex record: the 13 values of one
 * log_vortex_err.plt row after its time (src/mms.f90:351-361), then x, y of
 * the cell with the largest density error (log_vortex_err_xy.plt,
 * src/mms.f90:363). */
int fvs_time_integration(fvs_ctx *ctx, double t1, int nsteps, double *norms, double *verr);
/* No counterpart in the reference: fvs_time_integration for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same steps and outputs; the
 * per-step norms and vortex errors are staged in a fixed device array of the
 * library (one per GPU, held until the process ends) instead of buffers the
 * context grows, and steps are launched on the stream (never from a HIP
 * graph), so the call allocates nothing.  nsteps is at most 4096 (call it in
 * chunks); a larger nsteps fails with FVS_EINVAL.  Every context on one GPU
 * shares that staging: calls that overlap on one GPU (several threads or
 * contexts at once) race, and the caller must serialise them. */
int fvs_time_integration_noalloc(fvs_ctx *ctx, double t1, int nsteps, double *norms, double *verr);
/* error_isentropic_vortex (src/mms.f90:271-365) on the current state: out15
 * as one step's verr record of fvs_time_integration */
int fvs_vortex_error(fvs_ctx *ctx, double t, double *out15);
/* write_inst_cp_un C_L/C_D (src/io.f90:340-449), summed over slip/solid wall
 * boundaries: out4 = cl, cd, cl1, cd1 */
int fvs_wall_forces(fvs_ctx *ctx, double *out4);
/* The per-face part of write_inst_cp_un (src/io.f90:380-418): *nwall receives
 * the number of wall faces; out4 (may be NULL) receives per wall face [x_f,
 * cp with the face-extrapolated pressure, cp with the cell pressure, V_n with
 * the face-extrapolated velocity], faces in boundary-edge order (batch
 * contexts: instance by instance, each with its own Mach number); un3 (may
 * be NULL) receives max |V_n|, sqrt(mean V_n^2) and mean |V_n| over them
 * (the reference's figures for a mesh with one wall boundary). */
int fvs_wall_cp(fvs_ctx *ctx, int *nwall, double *out4, double *un3);
/* Change dt / CFL after creation (the reference reads them from fvs2d.input,
 * module input) */
int fvs_set_timestep(fvs_ctx *ctx, double dt, double cfl);
/* Measurement helper, no reference counterpart: runs nsteps steps like
 * fvs_time_integration (advancing the state) with HIP events around every
 * launch and returns, per kernel class [gradient(+limiter), interior-face
 * flux, boundary-face flux, gather+update, norm reduction, fused
 * flux+gather+update], the mean device
 * time per launch (ms), the algorithmic bytes per launch and the number of
 * launches timed. */
int fvs_kernel_times(fvs_ctx *ctx, double t1, int nsteps, double *ms6, double *bytes6, int *launches6);
/* Steady-state convergence aid, no reference counterpart: while frozen != 0
 * the limiter is not re-evaluated and every residual uses the limiter values
 * of the last evaluation. */
int fvs_freeze_limiter(fvs_ctx *ctx, int frozen);
/* Per-instance form of fvs_freeze_limiter for batch contexts, no reference
 * counterpart: frozen[k] != 0 keeps instance k's limiter values as last
 * evaluated while the other instances re-evaluate theirs every stage. */
int fvs_batch_freeze_limiter(fvs_ctx *ctx, const int *frozen);
/* Launch autotuner, no reference counterpart: times nsteps steps per trial
 * (each choice against the current setting, alternately, twice each, after
 * three warm-up trials) and keeps, per kernel class [gradient, face flux,
 * gather+update, fused flux+update], the faster of plain and XCD-aware
 * workgroup numbering, and the faster of 256 and 128 cells per workgroup for
 * the fused flux + update kernel and for the staged-halo LSQ gradient
 * (single meshes of at least 8,192 cells).  The state is restored
 * afterwards.  choice4 (may be NULL) receives the choices (bit 0 set =
 * XCD-aware; for the gradient, bit 1 set = the LSQ gradient, or a limited
 * Green-Gauss gradient's limiter, reads its stencil from a per-block halo
 * staged in LDS, the other candidate it tries, and bit 2 set = that LSQ
 * gradient on 128-cell workgroups; for the fused kernel, bit 2 set = 128-cell
 * workgroups), best_ms (may be NULL) the chosen setting's mean
 * time per step over its last two trials. */
int fvs_autotune(fvs_ctx *ctx, int nsteps, int *choice4, double *best_ms);
/* Applies launch choices in fvs_autotune's choice4 encoding (for instance
 * choices an earlier fvs_autotune returned), without timing anything; a
 * choice this context has no data for fails with FVS_EINVAL.  No reference
 * counterpart. */
int fvs_set_launch_choices(fvs_ctx *ctx, const int *choice4);
/* wall-clock of the last fvs_time_integration call's device work (ms) */
double fvs_last_elapsed_ms(const fvs_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
