"""Python mirror of the reference's module interface for the residual /
time-step path, over the C-ABI (include/fvs2d.h).

Method names follow the Fortran procedures they replace:

====================================  ===========================================
reference procedure                   here
====================================  ===========================================
grid_procs_init + gradient_init +     ``Solver(mesh, scheme)``
runge_kutta_init + data_solution_init
initialize_solution                   ``Solver.initialize_solution(mode, t0)``
cvar2pvar                             ``Solver.cvar2pvar()``
compute_gradient_cellcntr             ``Solver.compute_gradient_cellcntr()``
gradient_cellcntr_1var                ``Solver.gradient_cellcntr_1var(fc)``
compute_gradient_limiter              ``Solver.compute_gradient_limiter()``
compute_residual                      ``Solver.compute_residual(time)``
compute_flux_invscid                  :func:`compute_flux_invscid`
compute_local_time                    ``Solver.compute_local_time()``
time_integration                      ``Solver.time_integration(t1, n)``
error_isentropic_vortex               ``Solver.error_isentropic_vortex(t)``
write_inst_cp_un (C_L, C_D)           ``Solver.wall_forces()``
====================================  ===========================================

Arrays use the Fortran layouts transposed to C order: ``cvar``, ``pvar`` and
``resid`` are ``(ncells, 4)``, ``grad`` is ``(2, ncells, 4)``.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from typing import Optional, Tuple

import numpy as np

from . import _lib
from ._lib import FvsError, check, ptr
from .meshgen import Mesh


@dataclasses.dataclass
class Scheme:
    """Scheme flags with src/input.f90 semantics (see fvs2d_options)."""
    grad_method: int = 3      # 1 GGCB, 2 GGNB, 3 LSQ
    lsq_nn: int = 1           # 1 "nn", 0 "fn"
    lsq_p: float = 0.0
    limiter: int = 0          # 0 none, 1 venk, 2 barth, 3 albada
    recon: int = 2            # 1 upwind-1st, 2 upwind-2nd, 3 UMUSCL
    kappa: float = 0.0
    gamma: float = 1.4
    mach: float = 0.5
    aoa_deg: float = 0.0
    lvortex: int = 0
    vortex_x0: float = 0.0
    vortex_y0: float = 0.0
    vortex_K: float = 5.0
    vortex_inf: Tuple[float, float, float, float] = (1.0, 1.0, 1.0, 1.0)
    rk_scheme: int = _lib.RK_SSP3
    rk_order: int = 4
    steady: int = 0
    dt: float = 0.01
    cfl: float = 1.0
    reorder: int = 0          # device cell order: 0 reference, 1 Morton, 2 coordinate bisection
    kernel_path: int = 0      # 0 fastest available, 1 face-flux buffer, 2 whole stage, 3 no block halo

    def to_c(self) -> _lib.Options:
        o = _lib.Options()
        for f in dataclasses.fields(self):
            v = getattr(self, f.name)
            if f.name == "vortex_inf":
                o.vortex_inf = (C.c_double * 4)(*v)
            else:
                setattr(o, f.name, v)
        return o

    def oracle_kwargs(self) -> dict:
        """The scheme for the oracle, which evaluates SSP-RK(4,2) in the
        reference's form only: the two-register form (RK_SSP42_SO) maps to
        it (the same method, other rounding)."""
        d = dataclasses.asdict(self)
        d.pop("reorder")
        d.pop("kernel_path")
        if d["rk_scheme"] == _lib.RK_SSP42_SO:
            d["rk_scheme"] = _lib.RK_SSP42
        return d


def read_input(path: str) -> Tuple[str, Scheme, dict]:
    """Read a reference ``fvs2d.input`` (fixed line order, src/input.f90:86-120).

    Returns the grid base name, the scheme and the run control values
    (dt, ntimes, nsaves, ntstart)."""
    with open(path) as f:
        lines = f.read().splitlines()

    def val(i, k=0):
        return lines[i].replace(",", " ").split()[k]

    def flt(s):
        return float(s.lower().replace("d", "e"))

    def flag(s):
        return s.strip().upper().startswith("T")

    base = val(3).split("-")[0]
    sch = Scheme()
    sch.mach = flt(val(5))
    sch.aoa_deg = flt(val(6))
    sch.gamma = flt(val(7))
    run = dict(dt=flt(val(9)), ntimes=int(val(10)), nsaves=int(val(11)), ntstart=int(val(12)))
    sch.dt = run["dt"]
    sch.steady = int(flag(val(13, 0)))
    sch.cfl = flt(val(13, 1))
    sch.lvortex = int(flag(val(14)))
    if run["ntstart"] == 0:
        sch.lvortex = 0  # src/input.f90:140
    sch.grad_method = int(val(21, 0))
    sch.lsq_nn = int(val(21, 1).lower() == "nn")
    sch.lsq_p = flt(val(21, 2))
    sch.limiter = int(val(22))
    sch.recon = int(val(23, 0))
    sch.kappa = flt(val(23, 1)) if sch.recon == 3 else 0.0
    if int(val(24)) != 1:
        raise ValueError("only the Roe flux (1) exists")
    if int(val(26)) != 4:
        raise ValueError("the reference codes 4-stage Runge-Kutta only")
    sch.rk_order = int(val(27))
    ssp = flag(val(28))
    sch.rk_scheme = _lib.RK_SSP42 if (ssp and sch.rk_order == 2) else _lib.RK_CLASSIC
    return base, sch, run


def read_vortex(path: str, sch: Scheme) -> Scheme:
    """Read ``fvs2d.vortex`` (src/mms.f90:55-60) into the scheme."""
    with open(path) as f:
        v = [ln.replace(",", " ").split() for ln in f.read().splitlines()]
    flt = lambda s: float(s.lower().replace("d", "e"))
    sch.vortex_x0, sch.vortex_y0 = flt(v[0][0]), flt(v[0][1])
    sch.vortex_K = flt(v[1][0])
    sch.vortex_inf = tuple(flt(v[k][0]) for k in range(2, 6))
    return sch


def _mesh_arrays(mesh: Mesh):
    nb, bt, bc = mesh.bc_arrays()
    keep = [np.ascontiguousarray(mesh.xy, np.float64), np.ascontiguousarray(mesh.cell_nodes, np.int32),
            np.ascontiguousarray(nb), np.ascontiguousarray(bt), np.ascontiguousarray(bc)]
    canon = None if mesh.canon is None else np.ascontiguousarray(mesh.canon, np.int32)
    keep.append(canon)
    return keep


class HostMesh:
    """Host-only topology / stencils (fvs_mesh_*), no device needed."""

    def __init__(self, mesh: Mesh, grad_method=3, lsq_nn=1, lsq_p=0.0, limiter=0):
        L = _lib.load()
        xy, cn, nb, bt, bc, canon = _mesh_arrays(mesh)
        h = C.c_void_p()
        check(L.fvs_mesh_create(C.byref(h), mesh.nnodes, ptr(xy), mesh.ntri, mesh.nquad, ptr(cn), ptr(canon),
                                len(nb), ptr(nb), ptr(bt), ptr(bc), grad_method, lsq_nn, lsq_p, limiter))
        self._h, self._owned = h, True
        self._L = L
        self._load_sizes()

    @staticmethod
    def failed_instances(err: Exception):
        """Indices of the instances a batch preparation rejected, from its
        error message (empty when the error names none)."""
        import re
        m = re.search(r"instances \[([0-9,]+)\] failed", str(err))
        return [int(v) for v in m.group(1).split(",")] if m else []

    @classmethod
    def batch(cls, meshes, grad_method=3, lsq_nn=1, lsq_p=0.0, limiter=0, nthreads=0):
        """Host preparation of a batch of meshes, instance by instance on
        nthreads threads (fvs_mesh_create_batch), numbered as the
        concatenation of the meshes (meshgen.concat)."""
        from .meshgen import concat
        mesh, coff, _ = concat(meshes)
        L = _lib.load()
        xy, cn, nb, bt, bc, _ = _mesh_arrays(mesh)
        off = np.ascontiguousarray(coff, np.int32)
        h = C.c_void_p()
        check(L.fvs_mesh_create_batch(C.byref(h), len(meshes), ptr(off), mesh.nnodes, ptr(xy), mesh.ntri, mesh.nquad,
                                      ptr(cn), len(nb), ptr(nb), ptr(bt), ptr(bc), grad_method, lsq_nn, lsq_p,
                                      limiter, nthreads))
        obj = cls.__new__(cls)
        obj._h, obj._owned, obj._L = h, True, L
        obj._load_sizes()
        return obj

    @classmethod
    def _view(cls, L, h):
        obj = cls.__new__(cls)
        obj._h, obj._owned, obj._L = C.c_void_p(h), False, L
        obj._load_sizes()
        return obj

    def _load_sizes(self):
        s = np.zeros(8, np.int32)
        check(self._L.fvs_mesh_sizes(self._h, ptr(s)))
        (self.ncells, self.nedges, self.nnodes, self.nedges_intr, self.nedges_bndr, self.ncells_intr,
         self.ncells_bndr, self.nb) = (int(v) for v in s)

    def __del__(self):
        if getattr(self, "_owned", False) and self._h:
            self._L.fvs_mesh_destroy(self._h)
            self._h = None

    def edges(self):
        E = self.nedges
        n1, n2, c1, c2 = (np.zeros(E, np.int32) for _ in range(4))
        g = np.zeros(7 * E)
        check(self._L.fvs_mesh_topology(self._h, ptr(n1), ptr(n2), ptr(c1), ptr(c2), ptr(g), None, None, None))
        g = g.reshape(7, E)
        return dict(n1=n1, n2=n2, c1=c1, c2=c2, x=g[0], y=g[1], area=g[2], nx=g[3], ny=g[4], sx=g[5], sy=g[6])

    def cells(self):
        N = self.ncells
        xyv = np.zeros(3 * N)
        nghbre = np.zeros(4 * N, np.int32)
        edge = np.zeros(4 * N, np.int32)
        check(self._L.fvs_mesh_topology(self._h, None, None, None, None, None, ptr(xyv), ptr(nghbre), ptr(edge)))
        xyv = xyv.reshape(3, N)
        return dict(x=xyv[0], y=xyv[1], vol=xyv[2], nghbre=nghbre.reshape(N, 4), edge=edge.reshape(N, 4))

    def lists(self):
        ei = np.zeros(max(self.nedges_intr, 1), np.int32)
        ci = np.zeros(max(self.ncells_intr, 1), np.int32)
        bn = np.zeros(max(self.nb, 1), np.int32)
        be = np.zeros(max(self.nedges_bndr, 1), np.int32)
        check(self._L.fvs_mesh_lists(self._h, ptr(ei), ptr(ci), ptr(bn), ptr(be)))
        return dict(edge_intr=ei[: self.nedges_intr], cell_intr=ci[: self.ncells_intr], bnd_nedges=bn[: self.nb],
                    bnd_edges=be[: self.nedges_bndr])

    def lsq(self):
        nnz = self._L.fvs_mesh_lsq_nnz(self._h)
        p = np.zeros(self.ncells + 1, np.int32)
        cell = np.zeros(max(nnz, 1), np.int32)
        w = np.zeros(max(nnz, 1))
        coef = np.zeros(max(2 * nnz, 2))
        check(self._L.fvs_mesh_lsq(self._h, ptr(p), ptr(cell), ptr(w), ptr(coef)))
        return dict(ptr=p, cell=cell[:nnz], w=w[:nnz], coef=coef[: 2 * nnz].reshape(-1, 2))

    def cell_order(self, reorder: int, cell_offsets=None) -> np.ndarray:
        """Device cell order a solver on this mesh uses (fvs_mesh_cell_order):
        perm[i] = reference id of the cell at device position i; pass a batch's
        cell_offsets to keep its instances contiguous."""
        perm = np.zeros(self.ncells, np.int32)
        if cell_offsets is None:
            check(self._L.fvs_mesh_cell_order(self._h, reorder, 0, None, ptr(perm)))
        else:
            off = np.ascontiguousarray(cell_offsets, np.int32)
            check(self._L.fvs_mesh_cell_order(self._h, reorder, len(off) - 1, ptr(off), ptr(perm)))
        return perm


class Solver:
    """One mesh + scheme on the GPU (an fvs_ctx)."""

    def __init__(self, mesh: Mesh, scheme: Optional[Scheme] = None, **overrides):
        self.scheme = dataclasses.replace(scheme or Scheme(), **overrides)
        L = _lib.load()
        self._L = L
        xy, cn, nb, bt, bc, canon = _mesh_arrays(mesh)
        self._opts = self.scheme.to_c()
        h = C.c_void_p()
        check(L.fvs_create(C.byref(h), mesh.nnodes, ptr(xy), mesh.ntri, mesh.nquad, ptr(cn), ptr(canon), len(nb),
                           ptr(nb), ptr(bt), ptr(bc), C.byref(self._opts)))
        self._h = h
        s = np.zeros(8, np.int32)
        check(L.fvs_sizes(h, ptr(s)))
        (self.ncells, self.nedges, self.nnodes, self.nedges_intr, self.nedges_bndr, self.ncells_intr,
         self.ncells_bndr, self.nb) = (int(v) for v in s)
        self.mesh = HostMesh._view(L, L.fvs_ctx_mesh(h))
        self.ws_nrml = None
        self.resid = None

    def close(self):
        if getattr(self, "_h", None):
            self._L.fvs_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    # -- state ------------------------------------------------------------
    @property
    def cvar(self) -> np.ndarray:
        out = np.zeros((self.ncells, 4))
        check(self._L.fvs_get_state(self._h, ptr(out)))
        return out

    @cvar.setter
    def cvar(self, v):
        a = np.ascontiguousarray(v, np.float64).reshape(self.ncells, 4)
        check(self._L.fvs_set_state(self._h, ptr(a)))

    def initialize_solution(self, mode: int = 1, t0: float = 0.0):
        check(self._L.fvs_initialize(self._h, mode, t0))

    # -- operators --------------------------------------------------------
    def cvar2pvar(self) -> np.ndarray:
        out = np.zeros((self.ncells, 4))
        check(self._L.fvs_primitives(self._h, ptr(out)))
        return out

    def compute_gradient_cellcntr(self) -> np.ndarray:
        out = np.zeros((2, self.ncells, 4))
        check(self._L.fvs_gradient(self._h, ptr(out)))
        return out

    def gradient_cellcntr_1var(self, fc) -> np.ndarray:
        fc = np.ascontiguousarray(fc, np.float64)
        out = np.zeros((2, self.ncells))
        check(self._L.fvs_gradient_1var(self._h, ptr(fc), ptr(out)))
        return out

    def compute_gradient_limiter(self) -> np.ndarray:
        out = np.zeros(self.ncells)
        check(self._L.fvs_limiter(self._h, ptr(out)))
        return out

    def compute_residual(self, time: float = 0.0):
        r = np.zeros((self.ncells, 4))
        w = np.zeros(self.ncells)
        check(self._L.fvs_compute_residual(self._h, time, ptr(r), ptr(w)))
        self.resid, self.ws_nrml = r, w
        return r, w

    def compute_local_time(self) -> np.ndarray:
        out = np.zeros(self.ncells)
        check(self._L.fvs_local_dt(self._h, ptr(out)))
        return out

    def time_integration(self, t1: float, nsteps: int, vortex_error: bool = False):
        norms = np.zeros((nsteps, 4))
        verr = np.zeros((nsteps, 15)) if vortex_error else None
        check(self._L.fvs_time_integration(self._h, t1, nsteps, ptr(norms), ptr(verr)))
        return (norms, verr) if vortex_error else norms

    def freeze_limiter(self, frozen: bool = True):
        check(self._L.fvs_freeze_limiter(self._h, int(frozen)))

    def set_graph(self, on: bool = True):
        """Replay each time step from a captured HIP graph (fvs_set_graph)."""
        check(self._L.fvs_set_graph(self._h, int(on)))

    def set_persistent(self, on: bool = True):
        """Run each time-integration call as one persistent cooperative launch
        (fvs_set_persistent; needs kernel_path 2)."""
        check(self._L.fvs_set_persistent(self._h, int(on)))

    def set_phase_timers(self, on: bool = True):
        """Accumulate device time per phase (fvs_set_phase_timers)."""
        check(self._L.fvs_set_phase_timers(self._h, int(on)))

    def phase_times(self) -> np.ndarray:
        """Seconds per phase [gradient, limiter, flux, Runge-Kutta] (fvs_phase_times)."""
        out = np.zeros(4)
        check(self._L.fvs_phase_times(self._h, ptr(out)))
        return out

    def steady_solve(self, max_steps: int, drop: float = 1e-8, chunk: int = 200, t1: float = 0.0,
                     freeze_at: Optional[float] = None, freeze_after: Optional[int] = None,
                     freeze_stall: Optional[Tuple[int, float]] = None,
                     give_up: Optional[Tuple[int, float]] = None):
        """Pseudo-time stepping until the density residual of
        src/runge_kutta.f90:396-398 has dropped by ``drop`` relative to the first
        step (the reference runs a fixed step count; the stopping rule is this
        library's).  Norms come back after every ``chunk`` steps; the state may
        run up to chunk-1 steps past the crossing.  With ``freeze_at`` the
        limiter is frozen once the drop passes that level, with
        ``freeze_after`` after that many steps (limiter chatter otherwise
        stalls convergence); with ``freeze_stall = (window, factor)`` once the
        norm at a chunk end is above ``factor`` times the norm ``window`` steps
        earlier (the same rule as steady_solve_batch).  With ``give_up =
        (window, factor)`` the solve stops unconverged once, after the limiter
        was frozen, the norm at a chunk end is above ``factor`` times the norm
        ``window`` steps earlier.  Returns a dict with the iteration of the
        crossing (or None), the norm history, the iteration at which the
        limiter was frozen, whether a non-finite norm (divergence) stopped the
        loop and whether the give-up rule did."""
        hist = []
        done = 0
        conv_at = None
        frozen_at = None
        diverged = False
        r1 = None
        ends, ends_f = [], []
        gave_up = False
        while done < max_steps:
            n = min(chunk, max_steps - done)
            if freeze_after is not None and frozen_at is None:
                if done >= freeze_after:
                    self.freeze_limiter(True)
                    frozen_at = done
                else:
                    n = min(n, freeze_after - done)
            norms = self.time_integration(t1, n)
            hist.append(norms)
            if r1 is None:
                r1 = float(norms[0, 0])
            if not np.all(np.isfinite(norms)):
                diverged = True
                done += n
                break
            hit = np.nonzero(norms[:, 0] <= drop * r1)[0]
            if hit.size:
                conv_at = done + int(hit[0]) + 1
                done += n
                break
            done += n
            ends.append((done, float(norms[-1, 0])))
            if frozen_at is not None:
                ends_f.append(ends[-1])
                if give_up is not None and _stalled(ends_f, give_up):
                    gave_up = True
                    break
            elif ((freeze_at is not None and norms[-1, 0] <= freeze_at * r1)
                  or (freeze_stall is not None and _stalled(ends, freeze_stall))):
                self.freeze_limiter(True)
                frozen_at = done
                ends_f = [ends[-1]]
        if frozen_at is not None:
            self.freeze_limiter(False)
        h = np.concatenate(hist) if hist else np.zeros((0, 4))
        return dict(iterations=conv_at, steps_run=done, history=h, diverged=diverged, limiter_frozen_at=frozen_at,
                    drop=float(h[-1, 0] / r1) if len(h) else None, gave_up=gave_up)

    def write_save(self, base: str, itime: int, imach: int = 2):
        """The reference's save file of the current state (``<base>.cd`` +
        ``.s8``; src/io.f90:96-113, 166-178), readable by its restart."""
        from . import restart
        restart.write_save(base, self.cvar, self.nnodes, itime, imach)

    def read_cont(self, base: str, imach: int = 2):
        """Set the state from the reference's restart files
        (src/initialize.f90:56-77)."""
        from . import restart
        self.cvar = restart.read_cont(base, self.ncells, self.nnodes, imach)

    def error_isentropic_vortex(self, t: float) -> np.ndarray:
        out = np.zeros(15)
        check(self._L.fvs_vortex_error(self._h, t, ptr(out)))
        return out

    def wall_forces(self) -> np.ndarray:
        out = np.zeros(4)
        check(self._L.fvs_wall_forces(self._h, ptr(out)))
        return out

    def wall_cp(self):
        """Per wall face x_f, cp, cp_cell and V_n, and the max / L2 / L1 of V_n
        (fvs_wall_cp; src/io.f90:380-418)."""
        n = np.zeros(1, np.int32)
        check(self._L.fvs_wall_cp(self._h, ptr(n), None, None))
        out = np.zeros((int(n[0]), 4))
        un3 = np.zeros(3)
        check(self._L.fvs_wall_cp(self._h, ptr(n), ptr(out), ptr(un3)))
        return dict(x=out[:, 0], cp=out[:, 1], cp_cell=out[:, 2], un=out[:, 3], un_max=un3[0], un_l2=un3[1],
                    un_l1=un3[2])

    def set_timestep(self, dt: float, cfl: Optional[float] = None):
        cfl = self.scheme.cfl if cfl is None else cfl
        check(self._L.fvs_set_timestep(self._h, dt, cfl))
        self.scheme.dt, self.scheme.cfl = dt, cfl

    KERNEL_CLASSES = ("gradient", "flux_interior", "flux_boundary", "gather_update", "norm_reduce",
                      "flux_update_fused")

    def kernel_times(self, t1: float, nsteps: int):
        """Per kernel class: (mean ms per launch, algorithmic bytes per launch,
        launches) over nsteps steps (the state advances)."""
        ms, by = np.zeros(6), np.zeros(6)
        n = np.zeros(6, np.int32)
        check(self._L.fvs_kernel_times(self._h, t1, nsteps, ptr(ms), ptr(by), ptr(n)))
        return {k: (float(ms[i]), float(by[i]), int(n[i])) for i, k in enumerate(self.KERNEL_CLASSES)}

    def autotune(self, nsteps: int = 10):
        """Pick plain or XCD-aware workgroup numbering per kernel class (bit
        0), for the LSQ gradient or a limited Green-Gauss gradient's limiter
        whether it reads its stencil from a per-block halo staged in LDS (bit
        1), and for the fused flux kernel and the staged-halo LSQ gradient 128-
        or 256-cell blocks (bit 2); the state is left as it was.  Returns
        (choices, best ms per step)."""
        ch = np.zeros(4, np.int32)
        best = np.zeros(1)
        check(self._L.fvs_autotune(self._h, nsteps, ptr(ch), ptr(best)))
        names = ("gradient", "flux", "update", "fused")
        return {n: int(c) for n, c in zip(names, ch)}, float(best[0])

    def set_launch_choices(self, choices: dict):
        """Apply launch choices in the encoding autotune() returns
        (fvs_set_launch_choices)."""
        ch = np.array([int(choices.get(n, 0)) for n in ("gradient", "flux", "update", "fused")], np.int32)
        check(self._L.fvs_set_launch_choices(self._h, ptr(ch)))

    @property
    def last_elapsed_ms(self) -> float:
        return float(self._L.fvs_last_elapsed_ms(self._h))


class BatchSolver(Solver):
    """Many independent meshes in one context (fvs_create_batch): every RK
    stage is one launch over all instances, each instance with its own free
    stream, residual norms, convergence and forces."""

    def __init__(self, meshes, mach_aoa, scheme: Optional[Scheme] = None, **overrides):
        from .meshgen import concat
        self.scheme = dataclasses.replace(scheme or Scheme(), **overrides)
        mesh, coff, _ = concat(meshes)
        self.ninst = len(meshes)
        self.cell_offsets = coff.astype(np.int32)
        self.mach_aoa = np.ascontiguousarray(np.asarray(mach_aoa, np.float64).reshape(self.ninst, 2))
        L = _lib.load()
        self._L = L
        xy, cn, nb, bt, bc, _ = _mesh_arrays(mesh)
        self._opts = self.scheme.to_c()
        h = C.c_void_p()
        check(L.fvs_create_batch(C.byref(h), self.ninst, ptr(self.cell_offsets), ptr(self.mach_aoa), mesh.nnodes,
                                 ptr(xy), mesh.ntri, mesh.nquad, ptr(cn), len(nb), ptr(nb), ptr(bt), ptr(bc),
                                 C.byref(self._opts)))
        self._h = h
        s = np.zeros(8, np.int32)
        check(L.fvs_sizes(h, ptr(s)))
        (self.ncells, self.nedges, self.nnodes, self.nedges_intr, self.nedges_bndr, self.ncells_intr,
         self.ncells_bndr, self.nb) = (int(v) for v in s)
        self.mesh = HostMesh._view(L, L.fvs_ctx_mesh(h))
        self.active = np.ones(self.ninst, np.int32)

    def initialize_freestream(self):
        """Each instance at its own free stream (src/initialize.f90:36-43)."""
        g = self.scheme.gamma
        cv = np.zeros((self.ncells, 4))
        for k in range(self.ninst):
            M, a = self.mach_aoa[k]
            u, v = M * np.cos(np.radians(a)), M * np.sin(np.radians(a))
            p = 1.0 / g
            cv[self.cell_offsets[k]:self.cell_offsets[k + 1]] = [1.0, u, v, p / (g - 1) + 0.5 * (u * u + v * v)]
        self.cvar = cv

    def instance_state(self, k: int) -> np.ndarray:
        return self.cvar[self.cell_offsets[k]:self.cell_offsets[k + 1]]

    def batch_time_integration(self, t1: float, nsteps: int) -> np.ndarray:
        out = np.zeros((nsteps, self.ninst, 4))
        check(self._L.fvs_batch_time_integration(self._h, t1, nsteps, ptr(out)))
        return out

    def set_active(self, flags):
        self.active = np.ascontiguousarray(flags, np.int32)
        check(self._L.fvs_batch_set_active(self._h, ptr(self.active)))

    def wall_forces_batch(self) -> np.ndarray:
        out = np.zeros((self.ninst, 4))
        check(self._L.fvs_batch_wall_forces(self._h, ptr(out)))
        return out

    def freeze_limiter_batch(self, flags):
        """Freeze the limiter of the instances with flags[k] != 0 only
        (fvs_batch_freeze_limiter)."""
        f = np.ascontiguousarray(flags, np.int32)
        check(self._L.fvs_batch_freeze_limiter(self._h, ptr(f)))

    def freestream_state(self, k: int) -> np.ndarray:
        """Conservative free-stream state of instance k (src/initialize.f90:36-43)."""
        g = self.scheme.gamma
        M, a = self.mach_aoa[k]
        u, v = M * np.cos(np.radians(a)), M * np.sin(np.radians(a))
        return np.array([1.0, u, v, (1.0 / g) / (g - 1) + 0.5 * (u * u + v * v)])

    def reference_norms(self) -> np.ndarray:
        """Density-residual norm of one step from each instance's free stream,
        i.e. the first-step norm of a cold start.  The state is restored."""
        save = self.cvar
        self.initialize_freestream()
        r = self.batch_time_integration(0.0, 1)[0, :, 0].copy()
        self.cvar = save
        return r

    def steady_solve_batch(self, max_steps: int, drop: float = 1e-8, chunk: int = 200,
                           freeze_after: Optional[int] = None, freeze_drop: Optional[float] = None,
                           r_ref=None, diverge: Optional[float] = None, on_diverge=None,
                           freeze_stall: Optional[Tuple[int, float]] = None, progress=None,
                           max_active: Optional[int] = None, give_up: Optional[Tuple[int, float]] = None):
        """Pseudo-time stepping of all instances, each until its density
        residual norm is at most ``drop * r_ref[k]`` (``r_ref`` defaults to
        each instance's first-step norm); converged instances are masked out
        at the next chunk boundary.  Instance k's limiter is frozen once its
        norm is at most ``freeze_drop * r_ref[k]`` or after ``freeze_after`` of
        its own steps since its (re)start, or, with ``freeze_stall = (window,
        factor)``, once its norm at a chunk end is above ``factor`` times its
        norm ``window`` steps earlier (it stalls); ``window`` may be one value
        per instance.  It has diverged when a norm is
        non-finite or above ``diverge * r_ref[k]``; ``on_diverge(k)`` may then
        reset its state and return True to restart it, its step count running
        on, otherwise it is masked out.  Returns per-instance iterations (-1
        when not converged), steps run, divergence and restart flags, and the
        step at which each limiter was frozen (-1: never).  ``progress(steps,
        active)`` (optional) is called after every chunk.

        With ``max_active`` at most that many instances step at once: the
        others wait, in index order, with their initial state (masked blocks
        are not launched, so they cost nothing), and one starts at each chunk
        boundary for each instance that has finished (converged or given up).
        An instance's iterations, freezing and divergence count from its
        start; ``max_steps`` bounds each instance's own steps.

        With ``give_up = (window, factor)`` (``window`` may be one value per
        instance) an instance is given up, unconverged, once after its limiter
        was frozen its norm at a chunk end is above ``factor`` times its norm
        ``window`` steps earlier; ``gave_up`` flags them."""
        n = self.ninst
        cap = n if max_active is None else max(1, min(int(max_active), n))
        if freeze_stall is not None:
            win = np.broadcast_to(np.asarray(freeze_stall[0], np.int64), (n,))

            def stall_rule(k):
                return int(win[k]), freeze_stall[1]
        if give_up is not None:
            gwin = np.broadcast_to(np.asarray(give_up[0], np.int64), (n,))
        act0 = np.zeros(n, np.int32)
        act0[:cap] = 1
        self.set_active(act0)
        queued = list(range(cap, n))
        started = np.full(n, -1, np.int64)
        started[:cap] = 0
        iters = np.full(n, -1, np.int64)
        if r_ref is None and cap < n:
            raise ValueError("max_active below the instance count needs r_ref (each instance's reference norm)")
        r_ref = None if r_ref is None else np.asarray(r_ref, np.float64)
        done = 0
        since = np.zeros(n, np.int64)
        frozen = np.zeros(n, np.int32)
        frozen_at = np.full(n, -1, np.int64)
        diverged = np.zeros(n, bool)
        restarted = np.zeros(n, bool)
        gave_up = np.zeros(n, bool)
        ends = [[] for _ in range(n)]
        ends_f = [[] for _ in range(n)]
        self.freeze_limiter_batch(frozen)
        while np.any(self.active):
            m = chunk
            own = [max_steps - (done - started[k]) for k in np.nonzero(self.active)[0]]
            if min(own) <= 0:
                for k in np.nonzero(self.active)[0]:
                    if done - started[k] >= max_steps:
                        self.active[k] = 0
                while queued and int(np.sum(self.active)) < cap:
                    k = queued.pop(0)
                    self.active[k] = 1
                    started[k] = done
                if not np.any(self.active):
                    break
                self.set_active(self.active)
                continue
            m = min(m, min(own))
            if freeze_after is not None:
                left = [freeze_after - since[k] for k in np.nonzero(self.active & (frozen == 0))[0]]
                left = [v for v in left if v > 0]
                if left:
                    m = min(m, min(left))
            nr = self.batch_time_integration(0.0, m)[:, :, 0]
            if r_ref is None:
                r_ref = nr[0].copy()
            reset = []
            for k in np.nonzero(self.active)[0]:
                bad = ~np.isfinite(nr[:, k])
                if diverge is not None:
                    bad |= nr[:, k] > diverge * r_ref[k]
                if np.any(bad):
                    if on_diverge is not None and not restarted[k] and on_diverge(k):
                        restarted[k] = True
                        reset.append(k)
                    else:
                        diverged[k] = True
                        self.active[k] = 0
                    continue
                hit = np.nonzero(nr[:, k] <= drop * r_ref[k])[0]
                if hit.size:
                    iters[k] = done - started[k] + int(hit[0]) + 1
                    self.active[k] = 0
                    continue
                since[k] += m
                ends[k].append((int(since[k]), float(nr[-1, k])))
                if frozen[k]:
                    ends_f[k].append(ends[k][-1])
                    if give_up is not None and _stalled(ends_f[k], (int(gwin[k]), give_up[1])):
                        gave_up[k] = True
                        self.active[k] = 0
                elif ((freeze_drop is not None and nr[-1, k] <= freeze_drop * r_ref[k])
                      or (freeze_after is not None and since[k] >= freeze_after)
                      or (freeze_stall is not None and _stalled(ends[k], stall_rule(k)))):
                    frozen[k] = 1
                    frozen_at[k] = done - started[k] + m
                    ends_f[k] = [ends[k][-1]]
            for k in reset:
                since[k] = 0
                frozen[k] = 0
                frozen_at[k] = -1
                ends[k] = []
                ends_f[k] = []
            done += m
            self.freeze_limiter_batch(frozen)
            while queued and int(np.sum(self.active)) < cap:
                k = queued.pop(0)
                self.active[k] = 1
                started[k] = done
            if np.any(self.active):
                self.set_active(self.active)
            if progress is not None:
                progress(done, int(np.sum(self.active)))
        self.freeze_limiter(False)
        self.set_active(np.ones(n, np.int32))
        return dict(iterations=iters, steps_run=done, diverged=diverged, restarted=restarted, frozen_at=frozen_at,
                    r_ref=r_ref, gave_up=gave_up)


def _stalled(ends, rule) -> bool:
    """ends: (step, norm) at chunk ends since the (re)start.  True when the
    latest norm is above factor times the norm at least window steps before."""
    window, factor = rule
    step, r = ends[-1]
    for s0, r0 in reversed(ends[:-1]):
        if step - s0 >= window:
            return r > factor * r0
    return False


def compute_flux_invscid(pL, pR, nx, ny, gamma=1.4):
    """Roe flux of n face states on the device (src/flux_invscid.f90:17-32)."""
    pL = np.ascontiguousarray(np.atleast_2d(pL), np.float64)
    pR = np.ascontiguousarray(np.atleast_2d(pR), np.float64)
    n = pL.shape[0]
    nxy = np.ascontiguousarray(np.stack([np.broadcast_to(nx, n), np.broadcast_to(ny, n)], 1), np.float64)
    f = np.zeros((n, 4))
    w = np.zeros(n)
    check(_lib.load().fvs_roe_flux_batch(n, gamma, ptr(pL), ptr(pR), ptr(nxy), ptr(f), ptr(w)))
    return f, w
