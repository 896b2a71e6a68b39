"""ctypes wrapper of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module: as the reference results the tests compare with, or as the timed
CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libfvs_oracle.so")

RK_CLASSIC, RK_SSP42, RK_SSP3 = 0, 1, 2


class Opts(C.Structure):
    _fields_ = [
        ("grad_method", C.c_int), ("lsq_nn", C.c_int), ("lsq_p", C.c_double),
        ("limiter", C.c_int), ("recon", C.c_int), ("kappa", C.c_double),
        ("gamma", C.c_double), ("mach", C.c_double), ("aoa_deg", C.c_double),
        ("lvortex", C.c_int), ("vortex_x0", C.c_double), ("vortex_y0", C.c_double),
        ("vortex_K", C.c_double), ("vortex_inf", C.c_double * 4),
        ("rk_scheme", C.c_int), ("rk_order", C.c_int), ("steady", C.c_int),
        ("dt", C.c_double), ("cfl", C.c_double),
    ]


REF_SRC = "/root/reference/src/ios_unstrc.f90"
IOS_REF = os.path.join(HERE, "_ref", "ios_ref")
GAMBIT_REF = os.path.join(HERE, "_ref", "read_gambit")
FLANG = "/opt/rocm/bin/amdflang"


def build() -> str:
    """Compile the oracle library in place (gcc, via oracle/Makefile), and,
    where the reference sources and amdflang are present, the reference's
    restart-file module under its test driver (oracle/_ref/ios_ref) and its
    Gambit neutral-file converter (oracle/_ref/read_gambit)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if os.path.exists(REF_SRC) and os.path.exists(FLANG):
        subprocess.run(["make", "-s", "-B", "-C", HERE, "ref"], check=True)
    return LIB


def build_variant(mutant: int, outdir: str) -> str:
    """Compile the oracle with planted mistake `mutant` (the MUT_* switches of
    fvs_oracle.c) into outdir; for the test that shows the pins catch them."""
    out = os.path.join(outdir, f"libfvs_oracle_mut{mutant}.so")
    cmd = ["gcc", "-O1", "-fPIC", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-shared",
           f"-DORC_MUTANT={int(mutant)}", "-o", out, os.path.join(HERE, "fvs_oracle.c"), "-lm", "-fopenmp"]
    subprocess.run(cmd, check=True)
    return out


_libs = {}


def lib(path: str = LIB):
    if path not in _libs:
        if path == LIB and not os.path.exists(LIB):
            build()
        L = C.CDLL(path)
        P = C.c_void_p
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        L.orc_create.restype = P
        L.orc_create.argtypes = [C.c_int, dp, C.c_int, C.c_int, ip, C.c_void_p, C.c_int, ip, ip, ip,
                                 C.POINTER(Opts)]
        L.orc_last_error.restype = C.c_char_p
        L.orc_destroy.argtypes = [P]
        for name in ["orc_cvar2pvar", "orc_pvar2cvar", "orc_gradient", "orc_limiter", "orc_local_time"]:
            getattr(L, name).argtypes = [P]
        L.orc_compute_residual.argtypes = [P, C.c_double]
        L.orc_initialize.argtypes = [P, C.c_int, C.c_double]
        L.orc_time_integration.argtypes = [P, C.c_double, C.c_int, C.c_void_p, C.c_void_p]
        L.orc_set_threads.argtypes = [P, C.c_int]
        L.orc_roe.argtypes = [C.c_double, dp, dp, C.c_double, C.c_double, dp, dp]
        L.orc_limiter_fn.restype = C.c_double
        L.orc_limiter_fn.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
        L.orc_vortex_state.argtypes = [P, C.c_double, C.c_double, C.c_double, dp]
        L.orc_mms_state.argtypes = [P, C.c_double, C.c_double, dp, dp]
        L.orc_bc_state.argtypes = [P, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                   dp, dp]
        L.orc_vortex_error.argtypes = [P, C.c_double, dp]
        L.orc_wall_forces.argtypes = [P, dp]
        L.orc_wall_cp.argtypes = [P, dp, dp]
        L.orc_wall_cp.restype = C.c_int
        L.orc_mms_source.argtypes = [P, dp]
        L.orc_gradient_1var.argtypes = [P, dp, dp]
        L.orc_lsq_nnz.restype = C.c_int
        L.orc_lsq_nnz.argtypes = [P]
        _libs[path] = L
    return _libs[path]


def make_opts(**kw) -> Opts:
    o = Opts()
    d = dict(grad_method=3, lsq_nn=1, lsq_p=0.0, limiter=0, recon=2, kappa=0.0, gamma=1.4, mach=0.5,
             aoa_deg=0.0, lvortex=0, vortex_x0=0.0, vortex_y0=0.0, vortex_K=5.0,
             vortex_inf=(1.0, 1.0, 1.0, 1.0), rk_scheme=RK_SSP3, rk_order=4, steady=0, dt=0.01, cfl=1.0)
    d.update(kw)
    for k, v in d.items():
        if k == "vortex_inf":
            o.vortex_inf = (C.c_double * 4)(*v)
        else:
            setattr(o, k, v)
    return o


class Oracle:
    """One oracle instance over a mesh (fvs2d_amd.meshgen.Mesh); lib_path
    selects another build of the oracle (a planted-mistake variant)."""

    def __init__(self, mesh, lib_path: str = LIB, **opts):
        L = lib(lib_path)
        self.L = L
        self.opts = make_opts(**opts)
        self.mesh = mesh
        nb, bt, bc = mesh.bc_arrays()
        canon = None
        if mesh.canon is not None:
            self._canon = np.ascontiguousarray(mesh.canon, np.int32)
            canon = self._canon.ctypes.data
        self.h = L.orc_create(mesh.nnodes, np.ascontiguousarray(mesh.xy, np.float64), mesh.ntri, mesh.nquad,
                              np.ascontiguousarray(mesh.cell_nodes, np.int32), canon, len(nb),
                              np.ascontiguousarray(nb), np.ascontiguousarray(bt), np.ascontiguousarray(bc),
                              C.byref(self.opts))
        if not self.h:
            raise RuntimeError("oracle: " + L.orc_last_error().decode())
        s = np.zeros(8, np.int32)
        L.orc_sizes(C.c_void_p(self.h), s.ctypes.data_as(C.c_void_p))
        (self.ncells, self.nedges, self.nnodes, self.nedges_intr, self.nedges_bndr,
         self.ncells_intr, self.ncells_bndr, self.nb) = (int(v) for v in s)

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.L.orc_destroy(C.c_void_p(h))
            self.h = None

    # -- helpers -----------------------------------------------------------
    def _call(self, name, *arrays):
        getattr(self.L, name)(C.c_void_p(self.h), *[a.ctypes.data_as(C.c_void_p) for a in arrays])

    def _get(self, name, shape, dtype=np.float64):
        a = np.zeros(shape, dtype)
        self._call(name, a)
        return a

    # -- topology ----------------------------------------------------------
    def edges(self):
        E = self.nedges
        n1, n2, c1, c2 = (np.zeros(E, np.int32) for _ in range(4))
        g = np.zeros(7 * E)
        self._call("orc_get_edges", n1, n2, c1, c2, g)
        g = g.reshape(7, E)
        return dict(n1=n1, n2=n2, c1=c1, c2=c2, x=g[0], y=g[1], area=g[2], nx=g[3], ny=g[4], sx=g[5], sy=g[6])

    def cells(self):
        N = self.ncells
        xyv = np.zeros(3 * N)
        nvrt = np.zeros(N, np.int32)
        nghbre = np.zeros(4 * N, np.int32)
        edge = np.zeros(4 * N, np.int32)
        sign = np.zeros(4 * N)
        self._call("orc_get_cells", xyv, nvrt, nghbre, edge, sign)
        xyv = xyv.reshape(3, N)
        return dict(x=xyv[0], y=xyv[1], vol=xyv[2], nvrt=nvrt, nghbre=nghbre.reshape(N, 4),
                    edge=edge.reshape(N, 4), nrmlsign=sign.reshape(N, 4))

    def lists(self):
        ei = np.zeros(self.nedges_intr, np.int32)
        ci = np.zeros(self.ncells_intr, np.int32)
        bn = np.zeros(max(self.nb, 1), np.int32)
        be = np.zeros(max(self.nedges_bndr, 1), np.int32)
        self._call("orc_get_lists", ei, ci, bn, be)
        return dict(edge_intr=ei, cell_intr=ci, bnd_nedges=bn[: self.nb], bnd_edges=be[: self.nedges_bndr])

    def lsq(self):
        nnz = self.L.orc_lsq_nnz(C.c_void_p(self.h))
        ptr = np.zeros(self.ncells + 1, np.int32)
        cell = np.zeros(max(nnz, 1), np.int32)
        w = np.zeros(max(nnz, 1))
        coef = np.zeros(max(2 * nnz, 2))
        self._call("orc_get_lsq", ptr, cell, w, coef)
        return dict(ptr=ptr, cell=cell[:nnz], w=w[:nnz], coef=coef[: 2 * nnz].reshape(-1, 2))

    # -- state ---------------------------------------------------------------
    @property
    def cvar(self):
        return self._get("orc_get_cvar", (self.ncells, 4))

    @cvar.setter
    def cvar(self, v):
        self._call("orc_set_cvar", np.ascontiguousarray(v, np.float64))

    @property
    def pvar(self):
        return self._get("orc_get_pvar", (self.ncells, 4))

    @pvar.setter
    def pvar(self, v):
        self._call("orc_set_pvar", np.ascontiguousarray(v, np.float64))

    @property
    def grad(self):
        return self._get("orc_get_grad", (2, self.ncells, 4))

    @property
    def phi(self):
        return self._get("orc_get_phi", (self.ncells,))

    @property
    def resid(self):
        return self._get("orc_get_resid", (self.ncells, 4))

    @property
    def ws(self):
        return self._get("orc_get_ws", (self.ncells,))

    @property
    def dt_local(self):
        return self._get("orc_get_dt_local", (self.ncells,))

    # -- operators -------------------------------------------------------------
    def set_threads(self, n: int):
        """Per-cell work of later calls on n OpenMP threads (bit-identical to
        the serial evaluation; 1, the default, is serial)."""
        if self.L.orc_set_threads(C.c_void_p(self.h), int(n)) != 0:
            raise ValueError("thread count must be positive")

    def initialize(self, mode=1, t0=0.0):
        self.L.orc_initialize(C.c_void_p(self.h), mode, t0)

    def cvar2pvar(self):
        self.L.orc_cvar2pvar(C.c_void_p(self.h))

    def gradient(self):
        self.L.orc_gradient(C.c_void_p(self.h))

    def limiter(self):
        self.L.orc_limiter(C.c_void_p(self.h))

    def compute_residual(self, t=0.0):
        self.L.orc_compute_residual(C.c_void_p(self.h), t)

    def local_time(self):
        self.L.orc_local_time(C.c_void_p(self.h))

    def gradient_1var(self, fc):
        out = np.zeros((2, self.ncells))
        self.L.orc_gradient_1var(C.c_void_p(self.h), np.ascontiguousarray(fc, np.float64), out)
        return out

    def time_integration(self, t1, nsteps, want_verr=False):
        norms = np.zeros((nsteps, 4))
        verr = np.zeros((nsteps, 15)) if want_verr else None
        self.L.orc_time_integration(C.c_void_p(self.h), t1, nsteps, norms.ctypes.data,
                                    verr.ctypes.data if want_verr else None)
        return (norms, verr) if want_verr else norms

    def vortex_state(self, t, x, y):
        out = np.zeros(4)
        self.L.orc_vortex_state(C.c_void_p(self.h), t, x, y, out)
        return out

    def mms_state(self, x, y):
        sol, rhs = np.zeros(4), np.zeros(4)
        self.L.orc_mms_state(C.c_void_p(self.h), x, y, sol, rhs)
        return sol, rhs

    def bc_state(self, bc_type, t, x, y, nx, ny, pfl):
        out = np.zeros(4)
        rc = self.L.orc_bc_state(C.c_void_p(self.h), bc_type, t, x, y, nx, ny,
                                 np.ascontiguousarray(pfl, np.float64), out)
        if rc:
            raise RuntimeError(self.L.orc_last_error().decode())
        return out

    def vortex_error(self, t):
        out = np.zeros(15)
        self.L.orc_vortex_error(C.c_void_p(self.h), t, out)
        return out

    def wall_forces(self):
        out = np.zeros(4)
        self.L.orc_wall_forces(C.c_void_p(self.h), out)
        return out

    def wall_cp(self):
        """Per wall face [x_f, cp, cp_cell, V_n] and [max, L2, L1] of V_n
        (src/io.f90:380-418)."""
        ls = self.lists()
        nmax = int(np.sum(ls["bnd_nedges"]))
        out = np.zeros(4 * max(nmax, 1))
        un3 = np.zeros(3)
        k = self.L.orc_wall_cp(C.c_void_p(self.h), out, un3)
        return out[: 4 * k].reshape(k, 4), un3

    def mms_source(self):
        out = np.zeros((self.ncells, 4))
        self.L.orc_mms_source(C.c_void_p(self.h), out)
        return out


def roe(gamma, pl, pr, nx, ny):
    f, w = np.zeros(4), np.zeros(1)
    lib().orc_roe(gamma, np.ascontiguousarray(pl, np.float64), np.ascontiguousarray(pr, np.float64), nx, ny, f, w)
    return f, w[0]


def limiter_fn(kind, a, b, vol):
    return lib().orc_limiter_fn(kind, a, b, vol)
