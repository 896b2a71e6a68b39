"""The C-ABI library: loads without a GPU, exports every entry point that
include/fvs2d.h declares, rejects bad arguments with an error code and a
message, and fails loudly (no CPU fallback) when no device is present."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from fvs2d_amd import _lib, meshgen
from fvs2d_amd.solver import Scheme, Solver, read_input, read_vortex

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "fvs2d.h")
HDR_MLP = os.path.join(ROOT, "include", "fvs2d_mlp.h")
GOLD = os.path.join(ROOT, "tests", "golden")


def declared(path=HDR):
    src = open(path).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fvs_\w+)\s*\(", src)))


def test_header_symbols_exported():
    L = _lib.load()
    names = declared()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), f"{n} declared in include/fvs2d.h but not exported"
    assert sorted(_lib.EXPORTS) == names
    from fvs2d_amd import mlp
    names = declared(HDR_MLP)
    for n in names:
        assert hasattr(L, n), f"{n} declared in include/fvs2d_mlp.h but not exported"
    assert sorted(mlp.MLP_EXPORTS) == names


def test_mlp_bad_arguments():
    from fvs2d_amd import mlp
    L = mlp._load()
    h = C.c_void_p()
    for depth, batch, word in ((0, 64, b"depth"), (9, 64, b"depth"), (2, 100, b"batch"), (2, 0, b"batch")):
        o = mlp.MLPOptions(depth, batch, 1e-3, 0.9, 0.999, 1e-8, 0)
        assert L.fvs_mlp_create(C.byref(h), C.byref(o)) == -1
        assert word in L.fvs_last_error()
    o = mlp.MLPOptions(2, 64, -1.0, 0.9, 0.999, 1e-8, 0)
    assert L.fvs_mlp_create(C.byref(h), C.byref(o)) == -1 and b"Adam" in L.fvs_last_error()
    assert L.fvs_mlp_train(None, 1, None) == -1


def test_bad_arguments_return_codes():
    L = _lib.load()
    h = C.c_void_p()
    opts = Scheme().to_c()
    rc = L.fvs_create(C.byref(h), 0, None, 0, 0, None, None, 0, None, None, None, C.byref(opts))
    assert rc == -1 and b"null" in L.fvs_last_error()
    assert L.fvs_time_integration(None, 0.0, 1, None, None) == -1
    assert L.fvs_set_timestep(None, 0.1, 1.0) == -1


def test_argument_checks_before_the_device():
    """Options, empty meshes and batch offsets are rejected with a message
    before any device call (so these run without a GPU)."""
    import numpy as np
    L = _lib.load()
    m = meshgen.jittered_square(4)
    xy = np.ascontiguousarray(m.xy, np.float64)
    cn = np.ascontiguousarray(m.cell_nodes, np.int32)
    nb_, ty, cells = m.bc_arrays()
    P = lambda a: a.ctypes.data  # noqa: E731
    h = C.c_void_p()

    def create(**kw):
        o = Scheme(**kw).to_c()
        return L.fvs_create(C.byref(h), m.nnodes, P(xy), m.ntri, 0, P(cn), None, len(nb_), P(nb_), P(ty),
                            P(cells), C.byref(o))

    for kw, msg in ((dict(recon=0), b"recon"), (dict(limiter=4), b"limiter"), (dict(rk_scheme=4), b"rk_scheme"),
                    (dict(rk_scheme=0, rk_order=5), b"rk_order"), (dict(gamma=1.0), b"gamma")):
        assert create(**kw) == -1 and msg in L.fvs_last_error(), kw
    o = Scheme().to_c()
    assert L.fvs_create(C.byref(h), m.nnodes, P(xy), 0, 0, P(cn), None, len(nb_), P(nb_), P(ty), P(cells),
                        C.byref(o)) == -1 and b"empty" in L.fvs_last_error()
    ma = np.zeros(4)
    for off, msg in (([0, 0, m.ntri], b"increase"), ([0, 10, m.ntri - 1], b"run from 0"),
                     ([0, m.ntri, m.ntri], b"increase")):
        off = np.array(off, np.int32)
        rc = L.fvs_create_batch(C.byref(h), 2, P(off), P(ma), m.nnodes, P(xy), m.ntri, 0, P(cn), len(nb_),
                                P(nb_), P(ty), P(cells), C.byref(o))
        assert rc == -1 and msg in L.fvs_last_error(), off


def _gpu_present():
    return os.path.exists("/dev/kfd")


@pytest.mark.skipif(_gpu_present(), reason="checks the behaviour without a device")
def test_no_device_fails_loudly():
    with pytest.raises(_lib.FvsError, match="no HIP device"):
        Solver(meshgen.jittered_square(4), Scheme())
    from fvs2d_amd.mlp import WarmStartNet
    with pytest.raises(_lib.FvsError, match="no HIP device"):
        WarmStartNet(depth=1, batch=64)


def test_input_decks_of_the_reference():
    """The reference's example fvs2d.input files parse to the flags
    src/input.f90:181-277 gives them."""
    base, s, run = read_input(os.path.join(GOLD, "vortex_example.fvs2d.input"))
    assert base == "vortex" and s.grad_method == 3 and s.lsq_nn == 0 and s.lsq_p == 0.0
    assert s.limiter == 0 and s.recon == 2 and s.kappa == 0.0 and s.rk_scheme == _lib.RK_CLASSIC
    assert s.rk_order == 4 and s.steady == 0 and s.lvortex == 1 and run["dt"] == 0.01 and run["ntimes"] == 4000
    s = read_vortex(os.path.join(GOLD, "vortex_example.fvs2d.vortex"), s)
    assert (s.vortex_x0, s.vortex_y0, s.vortex_K) == (5.0, 5.0, 1.0)
    assert s.vortex_inf == (1.0, 0.2, 0.0, 1.0)
    base, s, run = read_input(os.path.join(GOLD, "naca0012_example.fvs2d.input"))
    assert base == "naca0012_omesh" and s.lsq_nn == 1 and s.rk_scheme == _lib.RK_SSP42 and s.rk_order == 2
    assert s.steady == 1 and s.cfl == 1.25 and s.mach == 0.8 and s.lvortex == 0


def test_grid_roundtrip(tmp_path):
    m = meshgen.jittered_square(5, seed=2)
    base = str(tmp_path / "sq")
    meshgen.write_grid(m, base)
    r = meshgen.read_grid(base)
    assert np.array_equal(r.xy, m.xy) and np.array_equal(r.cell_nodes, m.cell_nodes)
    assert r.boundaries[0][0] == "dirichlet" and np.array_equal(r.boundaries[0][1], m.boundaries[0][1])
