"""Pins of the CPU oracle (test infrastructure) against what the mathematics
fixes, independently of the oracle's own formulas.  No golden vectors exist in
the reference (it has no tests or stored outputs), and the reference build is
not possible here (every hot-path module needs an MPI installation), so each
oracle function is pinned by closed forms, textbook restatements in a
different form, invariants or brute force:

  flux_invscid_roe    consistency, Roe matrix via the Euler Jacobian and an
                      eigen-decomposition (numpy), entropy fix, antisymmetry,
                      rotation covariance, supersonic upwinding
  gradient_lsq        exactness for random linear fields (fn, nn, p=0,1,2),
                      weighted least squares by numpy lstsq
  gradient_ggcb/ggnb  zero for constants, exact for linear fields on an
                      equilateral mesh, first-order convergence on a sine;
                      GGCB exact for linear fields on graded rectangles (face
                      midpoints on the centre segment at unequal distances);
                      both equal to their textbook forms computed from raw
                      node data on a jittered mesh (tests/pin_checks.py)
  limiter             Venkatakrishnan's published form, Barth's min(1,a/b),
                      constant-field branch, Van Albada special values; the
                      assembly (bounds, face extrapolation, min over faces and
                      variables) against a raw-data restatement
  compute_residual    free-stream preservation, discrete conservation, MMS
                      truncation-error convergence; slip-wall ghost by the
                      closed-form mirrored-state Roe flux
  runge_kutta         one step of each scheme equals its stability polynomial
                      applied through the linearised residual (classical
                      orders 1-4 from the stage structure of the table)

tests/test_oracle_mutants.py builds the oracle with one planted mistake at a
time and shows the checks of tests/pin_checks.py fail on each.
  mms / vortex        closed forms of the isentropic vortex
  wall forces         closed-form C_N / C_A of a linear pressure field
"""
import math

import numpy as np
import pytest

from fvs2d_amd import meshgen
from oracle.oracle import Oracle, limiter_fn, roe
from tests import pin_checks
from tests.common import vortex_scheme
from tests.golden_meshes import load_example

G = 1.4


def cons(p, g=G):
    r, u, v, P = p
    return np.array([r, r * u, r * v, P / (g - 1) + 0.5 * r * (u * u + v * v)])


def phys_flux(p, nx, ny, g=G):
    r, u, v, P = p
    un = u * nx + v * ny
    E = P / (g - 1) + 0.5 * r * (u * u + v * v)
    return np.array([r * un, r * un * u + P * nx, r * un * v + P * ny, un * (E + P)])


def euler_jacobian(u, v, H, nx, ny, g=G):
    """dF_n/dU of the 2-D Euler flux written with u, v, H (conservative U)."""
    q2 = u * u + v * v
    un = u * nx + v * ny
    g1 = g - 1
    dp = g1 * np.array([q2 / 2, -u, -v, 1.0])
    A = np.zeros((4, 4))
    A[0] = [0, nx, ny, 0]
    A[1] = np.array([-u * un, un + u * nx, u * ny, 0]) + nx * dp
    A[2] = np.array([-v * un, v * nx, un + v * ny, 0]) + ny * dp
    A[3] = [g1 * q2 / 2 * un - H * un, H * nx - g1 * u * un, H * ny - g1 * v * un, g * un]
    return A


def roe_textbook(pL, pR, nx, ny, fix=0.2, g=G):
    """0.5(F_L+F_R) - 0.5 R|Lambda|R^-1 (U_R-U_L) at the Roe average, with
    Harten's fix on the acoustic eigenvalues (Roe 1981; Harten 1983)."""
    rL, rR = pL[0], pR[0]
    HL = g / (g - 1) * pL[3] / rL + 0.5 * (pL[1] ** 2 + pL[2] ** 2)
    HR = g / (g - 1) * pR[3] / rR + 0.5 * (pR[1] ** 2 + pR[2] ** 2)
    sL, sR = math.sqrt(rL), math.sqrt(rR)
    u = (sL * pL[1] + sR * pR[1]) / (sL + sR)
    v = (sL * pL[2] + sR * pR[2]) / (sL + sR)
    H = (sL * HL + sR * HR) / (sL + sR)
    A = euler_jacobian(u, v, H, nx, ny, g)
    a = math.sqrt((g - 1) * (H - 0.5 * (u * u + v * v)))
    un = u * nx + v * ny
    # |A| by Sylvester's formula over the distinct eigenvalues un-a, un, un+a
    # (A is diagonalisable; un is a double eigenvalue)
    lam = [un - a, un, un + a]
    f = [abs(x) for x in lam]
    for k in (0, 2):  # Harten's fix on the acoustic waves
        if f[k] < fix:
            f[k] = 0.5 * (f[k] ** 2 / fix + fix)
    I = np.eye(4)
    D = np.zeros((4, 4))
    for k in range(3):
        P = I.copy()
        for j in range(3):
            if j != k:
                P = P @ (A - lam[j] * I) / (lam[k] - lam[j])
        D += f[k] * P
    F = 0.5 * (phys_flux(pL, nx, ny, g) + phys_flux(pR, nx, ny, g)) - 0.5 * D @ (cons(pR, g) - cons(pL, g))
    return F, A, (u, v, H, a)


def rand_state(rng, lo=0.3):
    return np.array([rng.uniform(lo, 2.0), rng.normal(0, 0.8), rng.normal(0, 0.8), rng.uniform(lo, 2.0)])


# ----------------------------------------------------------------------------- Roe
def test_roe_consistency():
    rng = np.random.default_rng(0)
    for _ in range(200):
        p = rand_state(rng)
        th = rng.uniform(0, 2 * np.pi)
        f, ws = roe(G, p, p, math.cos(th), math.sin(th))
        ex = phys_flux(p, math.cos(th), math.sin(th))
        assert np.allclose(f, ex, rtol=1e-14, atol=1e-14 * np.abs(ex).max())
        un = p[1] * math.cos(th) + p[2] * math.sin(th)
        assert math.isclose(ws, 0.5 * (abs(un) + math.sqrt(G * p[3] / p[0])), rel_tol=1e-14)


def test_roe_matches_jacobian_eigen_form():
    rng = np.random.default_rng(1)
    n = 0
    while n < 200:
        pL, pR = rand_state(rng), rand_state(rng)
        th = rng.uniform(0, 2 * np.pi)
        nx, ny = math.cos(th), math.sin(th)
        F, A, (u, v, H, a) = roe_textbook(pL, pR, nx, ny)
        # Roe's property: A(hat) dU = dF
        dF = phys_flux(pR, nx, ny) - phys_flux(pL, nx, ny)
        assert np.allclose(A @ (cons(pR) - cons(pL)), dF, rtol=1e-10, atol=1e-12 * np.abs(dF).max())
        f, ws = roe(G, pL, pR, nx, ny)
        assert np.allclose(f, F, rtol=1e-10, atol=1e-11 * max(1.0, np.abs(F).max()))
        assert math.isclose(ws, 0.5 * (abs(u * nx + v * ny) + a), rel_tol=1e-12)
        n += 1


def test_roe_entropy_fix_sonic():
    """Near-sonic Roe state: the acoustic |lambda| is replaced by
    (lambda^2/delta + delta)/2 with delta = 0.2."""
    rng = np.random.default_rng(2)
    hits = 0
    for _ in range(400):
        pL = rand_state(rng)
        a = math.sqrt(G * pL[3] / pL[0])
        pL[1], pL[2] = a + rng.uniform(-0.1, 0.1), 0.0
        pR = pL * (1 + 0.01 * rng.standard_normal(4))
        F, _, (u, v, H, ah) = roe_textbook(pL, pR, 1.0, 0.0)
        if abs(u - ah) < 0.2:
            hits += 1
        f, _ = roe(G, pL, pR, 1.0, 0.0)
        assert np.allclose(f, F, rtol=1e-9, atol=1e-10)
    assert hits > 100


def test_roe_antisymmetry_and_rotation():
    rng = np.random.default_rng(3)
    for _ in range(100):
        pL, pR = rand_state(rng), rand_state(rng)
        th = rng.uniform(0, 2 * np.pi)
        nx, ny = math.cos(th), math.sin(th)
        f1, w1 = roe(G, pL, pR, nx, ny)
        f2, w2 = roe(G, pR, pL, -nx, -ny)
        assert np.allclose(f1, -f2, rtol=1e-13, atol=1e-13 * np.abs(f1).max())
        assert math.isclose(w1, w2, rel_tol=1e-13)
        # rotate everything by phi: momentum flux rotates, scalars invariant
        ph = rng.uniform(0, 2 * np.pi)
        c, s = math.cos(ph), math.sin(ph)
        rot = lambda p: np.array([p[0], c * p[1] - s * p[2], s * p[1] + c * p[2], p[3]])
        f3, _ = roe(G, rot(pL), rot(pR), c * nx - s * ny, s * nx + c * ny)
        assert np.allclose(f3, rot(f1), rtol=1e-12, atol=1e-12 * np.abs(f1).max())


def test_roe_supersonic_upwinding():
    rng = np.random.default_rng(4)
    for _ in range(100):
        pL = rand_state(rng, lo=0.8)
        a = math.sqrt(G * pL[3] / pL[0])
        pL[1] = 2.5 * a + 0.5
        pR = pL * (1 + 0.02 * rng.standard_normal(4))
        f, _ = roe(G, pL, pR, 1.0, 0.0)
        ex = phys_flux(pL, 1.0, 0.0)
        assert np.allclose(f, ex, rtol=1e-12, atol=1e-12 * np.abs(ex).max())


# ----------------------------------------------------------------------------- gradients
def linear_field(o, a, b, c):
    cl = o.cells()
    return a + b * cl["x"] + c * cl["y"]


@pytest.mark.parametrize("nn", [0, 1])
@pytest.mark.parametrize("p", [0.0, 1.0, 2.0])
def test_lsq_linear_exact(nn, p):
    rng = np.random.default_rng(5)
    for m in (meshgen.jittered_square(10, seed=2), load_example("vortex_example")):
        o = Oracle(m, grad_method=3, lsq_nn=nn, lsq_p=p)
        a, b, c = rng.normal(size=3)
        g = o.gradient_1var(linear_field(o, a, b, c))
        assert np.allclose(g[0], b, atol=1e-10) and np.allclose(g[1], c, atol=1e-10)


def test_lsq_is_weighted_least_squares():
    """At each cell the gradient minimises sum_j w_j^2 (df_j - g.dx_j)^2 with
    w = 1/d^p (weights enter squared: rows of D are w*dx)."""
    m = meshgen.jittered_square(6, seed=4)
    rng = np.random.default_rng(6)
    for nn, p in ((1, 1.0), (0, 2.0)):
        o = Oracle(m, grad_method=3, lsq_nn=nn, lsq_p=p)
        cl, st = o.cells(), o.lsq()
        f = np.sin(cl["x"]) + rng.normal(size=o.ncells) * 0.1
        g = o.gradient_1var(f)
        for i in range(0, o.ncells, 7):
            js = st["cell"][st["ptr"][i]:st["ptr"][i + 1]]
            dx = np.stack([cl["x"][js] - cl["x"][i], cl["y"][js] - cl["y"][i]], 1)
            w = 1.0 / np.hypot(dx[:, 0], dx[:, 1]) ** p
            sol = np.linalg.lstsq(w[:, None] * dx, w * (f[js] - f[i]), rcond=None)[0]
            assert np.allclose(g[:, i], sol, rtol=1e-10, atol=1e-12)


def equilateral_periodic(nx=12, ny=12):
    """Periodic mesh of equilateral triangles (rows shifted by half a side)."""
    h = math.sqrt(3) / 2
    xy, canon = [], []
    for j in range(ny + 1):
        for i in range(nx + 1):
            xy.append((i + 0.5 * (j % 2), j * h))
            canon.append((j % ny) * (nx + 1) + (i % nx))
    idx = lambda i, j: j * (nx + 1) + i
    tri = []
    for j in range(ny):
        for i in range(nx):
            if j % 2 == 0:
                tri += [(idx(i, j), idx(i + 1, j), idx(i, j + 1)), (idx(i + 1, j), idx(i + 1, j + 1), idx(i, j + 1))]
            else:
                tri += [(idx(i, j), idx(i + 1, j + 1), idx(i, j + 1)), (idx(i, j), idx(i + 1, j), idx(i + 1, j + 1))]
    return meshgen.Mesh(np.array(xy), len(tri), 0, np.array(tri, np.int32).ravel(), [],
                        canon=np.array(canon, np.int32))


@pytest.mark.parametrize("gm", [1, 2])
def test_gg_zero_for_constant(gm):
    m = meshgen.jittered_square(10, seed=1)
    o = Oracle(m, grad_method=gm)
    g = o.gradient_1var(np.full(o.ncells, 3.7))
    assert np.max(np.abs(g)) < 1e-12


@pytest.mark.parametrize("gm", [1, 2])
def test_gg_linear_exact_on_equilateral_mesh(gm):
    """On equilateral triangles the face midpoint is the midpoint of the two
    centres (GGCB face value exact) and every node is the centre of its six
    cells' centroids (GGNB node value exact), so both are exact for linear
    fields away from the periodic seam."""
    m = equilateral_periodic()
    o = Oracle(m, grad_method=gm)
    cl = o.cells()
    f = 0.3 + 1.7 * cl["x"] - 0.6 * cl["y"]
    g = o.gradient_1var(f)
    far = (cl["x"] > 2.5) & (cl["x"] < 9.5) & (cl["y"] > 2.5) & (cl["y"] < 8)
    assert far.sum() > 50
    assert np.allclose(g[0][far], 1.7, atol=1e-12) and np.allclose(g[1][far], -0.6, atol=1e-12)


@pytest.mark.parametrize("gm,jitter,rate", [(3, 0.1, 0.95), (3, 0.0, 0.95), (1, 0.0, 0.95), (2, 0.0, 0.45)])
def test_gradient_converges(gm, jitter, rate):
    """Error of d/dx sin(2x)cos(3y) over interior cells under refinement.
    LSQ converges at first order on jittered meshes; the Green-Gauss forms are
    checked on the unjittered split-quad mesh (first order for GGCB, about
    half order for GGNB, whose inverse-distance node values are not exact for
    linear fields there): on jittered triangles both stall, the textbook
    inconsistency of Green-Gauss gradients on skewed cells."""
    errs = []
    for n in (16, 32, 64):
        m = meshgen.jittered_square(n, seed=0, lo=0.0, hi=1.0, jitter=jitter)
        o = Oracle(m, grad_method=gm)
        cl, ls = o.cells(), o.lists()
        f = np.sin(2 * cl["x"]) * np.cos(3 * cl["y"])
        g = o.gradient_1var(f)
        ex = 2 * np.cos(2 * cl["x"]) * np.cos(3 * cl["y"])
        ci = ls["cell_intr"]
        errs.append(np.sqrt(np.mean((g[0][ci] - ex[ci]) ** 2)))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(rates > rate), (errs, rates)


def test_ggcb_exact_on_graded_rectangles():
    pin_checks.check_ggcb_graded_exact()


@pytest.mark.parametrize("gm", [1, 2])
def test_gg_equals_raw_data_textbook_form(gm):
    (pin_checks.check_ggcb_brute if gm == 1 else pin_checks.check_ggnb_brute)()


# ----------------------------------------------------------------------------- limiter
@pytest.mark.parametrize("gm,lim", [(3, 1), (3, 2), (1, 1), (2, 2)])
def test_limiter_assembly_raw_data(gm, lim):
    pin_checks.check_limiter_brute(grad_method=gm, limiter=lim)


def test_slip_wall_closed_form():
    pin_checks.check_slip_wall_closed_form()


def test_venkatakrishnan_published_form():
    """phi = [(D+^2 + e^2) D- + 2 D-^2 D+] / [D+^2 + 2 D-^2 + D+ D- + e^2] / D-
    (Venkatakrishnan 1993), e^2 = (K h)^3, K = 5, h = 2 sqrt(V/pi)."""
    rng = np.random.default_rng(7)
    for _ in range(500):
        dp = rng.uniform(-1, 1)
        dm = rng.uniform(-1, 1)
        if abs(dm) < 1e-3:
            continue
        vol = rng.uniform(1e-4, 1e-1)
        e2 = (5.0 * 2.0 * math.sqrt(vol / math.pi)) ** 3
        ref = ((dp * dp + e2) * dm + 2 * dm * dm * dp) / (dp * dp + 2 * dm * dm + dp * dm + e2) / dm
        assert math.isclose(limiter_fn(1, dp, dm, vol), ref, rel_tol=1e-12, abs_tol=1e-14)


def test_barth_and_albada_closed_forms():
    assert limiter_fn(2, 0.5, 1.0, 1.0) == 0.5
    assert limiter_fn(2, 3.0, 1.0, 1.0) == 1.0
    assert limiter_fn(2, -0.2, -0.8, 1.0) == 0.25
    # the reference's Van Albada: ((b^2+e)a + (a^2+e)b)/(a^2+b^2+2e)/(b+e),
    # so at a == b it is b/(b+e) with e = (0.3 h)^3
    vol = 0.01
    e = (0.3 * 2.0 * math.sqrt(vol / math.pi)) ** 3
    for b in (0.1, 1.0, -0.3):
        assert math.isclose(limiter_fn(3, b, b, vol), b / (b + e), rel_tol=1e-14)


@pytest.mark.parametrize("lim", [1, 2, 3])
def test_limiter_constant_field_is_one(lim):
    m = meshgen.jittered_square(8, seed=1, periodic=True)
    o = Oracle(m, grad_method=3, limiter=lim)
    o.cvar = np.tile([1.0, 0.3, -0.2, 2.5], (o.ncells, 1))
    o.compute_residual(0.0)
    assert np.all(o.phi == 1.0)


@pytest.mark.parametrize("lim", [1, 2])
def test_limiter_bounded(lim):
    m = meshgen.jittered_square(16, seed=1)
    o = Oracle(m, **vortex_scheme(limiter=lim).oracle_kwargs())
    o.initialize(1, 0.0)
    o.compute_residual(0.0)
    assert np.all(o.phi <= 1.0) and np.all(o.phi >= 0.0)


# ----------------------------------------------------------------------------- residual
@pytest.mark.parametrize("gm,lim", [(3, 0), (3, 1), (1, 0), (2, 1)])
def test_freestream_preservation(gm, lim):
    # rho = p = 1: the vortex formula (src/mms.f90:236-253) assumes
    # rho_inf = p_inf = 1, so only then is a strength-0 vortex this state
    st = np.array([1.0, 0.5, 0.2, 1.0])
    for m in (meshgen.jittered_square(12, seed=2, periodic=True), load_example("vortex_example")):
        # Dirichlet boundaries carry a vortex of strength 0 = the same state
        o = Oracle(m, grad_method=gm, limiter=lim, lvortex=1, vortex_K=0.0,
                   vortex_inf=(st[0], st[1], st[2], st[3]))
        o.pvar = np.tile(st, (o.ncells, 1))
        o.L.orc_pvar2cvar(o.h)
        o.compute_residual(0.7)
        assert np.max(np.abs(o.resid)) < 1e-12


def test_discrete_conservation_periodic():
    m = meshgen.jittered_square(12, seed=2, periodic=True)
    o = Oracle(m, **vortex_scheme(limiter=1).oracle_kwargs())
    o.initialize(1, 0.0)
    o.compute_residual(0.0)
    vr = o.resid * o.cells()["vol"][:, None]
    assert np.all(np.abs(vr.sum(0)) <= 1e-13 * np.abs(vr).sum(0))


def test_mms_truncation_error():
    """test_resid (src/test.f90:481-559): resid + source over interior cells.
    Momentum and energy converge under refinement; continuity does not,
    because the reference's source term (src/mms.f90:169) has u*rx where the
    derivative of rho*u needs rho*ux."""
    e = []
    for n in (16, 32, 64):
        m = meshgen.jittered_square(n, seed=0, lo=0.0, hi=1.0)
        o = Oracle(m, grad_method=3, lsq_nn=1, lvortex=0, recon=2)
        o.initialize(0, 0.0)
        o.compute_residual(0.0)
        err = o.resid + o.mms_source()
        ci = o.lists()["cell_intr"]
        e.append(np.sqrt(np.mean(err[ci] ** 2, axis=0)))
    e = np.array(e)
    rates = np.log2(e[:-1] / e[1:])
    assert np.all(rates[:, 1:] > 0.8), rates
    assert np.all(np.abs(rates[:, 0]) < 0.3) and e[-1, 0] > 0.1, e


# ----------------------------------------------------------------------------- RK
def _poly_step(o, u0, coeffs, nfd=1e-4):
    """sum_k c_k (dt J)^k (u - u0) with J by central differences."""
    dt = o.opts.dt

    def R(u):
        o.cvar = u
        o.compute_residual(0.0)
        return o.resid.copy()

    def Jv(v):
        s = nfd / max(np.abs(v).max(), 1e-300)
        return (R(u0 + s * v) - R(u0 - s * v)) / (2 * s)

    return coeffs, dt, Jv


@pytest.mark.parametrize("scheme,coeffs", [
    (dict(rk_scheme=2), [1, 1, 1 / 2, 1 / 6, 0]),                       # SSP-RK3
    (dict(rk_scheme=0, rk_order=4), [1, 1, 1 / 2, 1 / 6, 1 / 24]),      # classical RK4
    (dict(rk_scheme=1, rk_order=2), [1, 1, 1 / 2, 1 / 9, 1 / 108]),     # SSP-RK(4,2)
])
def test_rk_stability_polynomial(scheme, coeffs):
    m = meshgen.jittered_square(8, seed=1, periodic=True)
    o = Oracle(m, **vortex_scheme(lvortex=0, dt=0.05, **scheme).oracle_kwargs())
    u0 = np.tile([1.0, 0.5, 0.2, 2.0], (o.ncells, 1))
    rng = np.random.default_rng(8)
    eps = 1e-6
    d = rng.standard_normal(u0.shape)
    _, dt, Jv = _poly_step(o, u0, coeffs)
    pred = np.zeros_like(d)
    v = d.copy()
    for k, c in enumerate(coeffs):
        pred += c * v
        if k < len(coeffs) - 1:
            v = dt * Jv(v)
    o.cvar = u0 + eps * d
    o.time_integration(0.0, 1)
    got = (o.cvar - u0) / eps
    assert np.max(np.abs(got - pred)) / np.max(np.abs(pred)) < 1e-4
    # the next polynomial coefficient matters: (dt J)^3 d is not negligible
    assert np.max(np.abs(v)) > 1e-3 * np.max(np.abs(pred))


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_classical_rk_tables(order):
    pin_checks.check_rk_polynomial(order=order)


# ----------------------------------------------------------------------------- vortex, forces
def test_isentropic_vortex_closed_form():
    m = meshgen.jittered_square(4)
    o = Oracle(m, **vortex_scheme(vortex_x0=1.0, vortex_y0=-0.5, vortex_K=5.0,
                                  vortex_inf=(1.0, 0.7, 0.3, 1.0)).oracle_kwargs())
    rng = np.random.default_rng(9)
    for _ in range(50):
        t, x, y = rng.uniform(0, 2), rng.uniform(-4, 4), rng.uniform(-4, 4)
        r, u, v, p = o.vortex_state(t, x, y)
        dx, dy = x - (1.0 + 0.7 * t), y - (-0.5 + 0.3 * t)
        rr = math.hypot(dx, dy)
        swirl = 5.0 / (2 * math.pi) * rr * math.exp(0.5 * (1 - rr * rr))
        # velocity perturbation is tangential with the closed-form magnitude
        du, dv = u - 0.7, v - 0.3
        assert math.isclose(math.hypot(du, dv), swirl, rel_tol=1e-12, abs_tol=1e-14)
        assert abs(du * dx + dv * dy) < 1e-12
        T = 1.0 - (G - 1) * 25.0 / (8 * G * math.pi ** 2) * math.exp(1 - rr * rr)
        assert math.isclose(p / r, T, rel_tol=1e-12)
        assert math.isclose(p, r ** G, rel_tol=1e-12)


def test_vortex_error_record_brute_force():
    """error_isentropic_vortex (src/mms.f90:271-365) restated in numpy from the
    closed-form vortex at the interior cell centres: max, mean and RMS of
    |U - U_exact| per conservative variable, the total RMS, and the centre of
    the first cell holding the largest density error (maxloc)."""
    m = meshgen.jittered_square(12, seed=3)
    o = Oracle(m, **vortex_scheme(vortex_x0=0.4, vortex_K=4.0).oracle_kwargs())
    o.initialize(1, 0.0)
    rng = np.random.default_rng(21)
    o.cvar = o.cvar * (1.0 + 1e-3 * rng.standard_normal((o.ncells, 4)))
    t = 0.3
    got = o.vortex_error(t)
    cl, ci = o.cells(), o.lists()["cell_intr"]
    d = np.zeros((ci.size, 4))
    cv = o.cvar
    for k, ic in enumerate(ci):
        r, u, v, p = o.vortex_state(t, cl["x"][ic], cl["y"][ic])
        ex = np.array([r, r * u, r * v, p / (G - 1) + 0.5 * r * (u * u + v * v)])
        d[k] = np.abs(cv[ic] - ex)
    n = ci.size
    ref = np.concatenate([np.stack([d.max(0), d.sum(0) / n, np.sqrt((d ** 2).sum(0) / n)], 1).ravel(),
                          [np.sqrt((d ** 2).sum() / n)]])
    assert np.allclose(got[:13], ref, rtol=1e-13, atol=0)
    erho = np.zeros(o.ncells)
    erho[ci] = d[:, 0]
    imax = int(np.argmax(erho))
    assert got[13] == cl["x"][imax] and got[14] == cl["y"][imax]


def test_wall_forces_linear_pressure():
    """p = a + b x + c y with exact (LSQ) gradients: the wall pressure is exact
    at each face midpoint, and the midpoint rule integrates a linear field
    exactly on the polygonal body, so with normals pointing into the body
    C_N = -(2/M^2) c A_body and C_A = -(2/M^2) b A_body."""
    m = load_example("naca0012_example")
    M = 0.8
    o = Oracle(m, grad_method=3, lsq_nn=1, mach=M, aoa_deg=1.25)
    cl = o.cells()
    b, c = 0.3, -0.7
    pv = np.zeros((o.ncells, 4))
    pv[:, 0] = 1.0
    pv[:, 3] = 2.0 + b * cl["x"] + c * cl["y"]
    o.pvar = pv
    out = o.wall_forces()
    # body polygon = wall edges' nodes
    e, ls = o.edges(), o.lists()
    we = ls["bnd_edges"][: ls["bnd_nedges"][0]]
    x1, y1 = m.xy[e["n1"][we], 0], m.xy[e["n1"][we], 1]
    x2, y2 = m.xy[e["n2"][we], 0], m.xy[e["n2"][we], 1]
    area = abs(0.5 * np.sum(x1 * y2 - x2 * y1))
    cn, ca = -2 / M ** 2 * c * area, -2 / M ** 2 * b * area
    al = math.radians(1.25)
    assert math.isclose(out[0], cn * math.cos(al) - ca * math.sin(al), rel_tol=1e-9)
    assert math.isclose(out[1], cn * math.sin(al) + ca * math.cos(al), rel_tol=1e-9)


def test_wall_cp_linear_fields():
    """Linear pressure and velocity fields with exact (LSQ) gradients: at each
    wall face midpoint cp = (2/M^2)(p(x_f) - p_inf) and V_n = (u(x_f),
    v(x_f)).n exactly; the cell-pressure cp uses the centroid value; and the
    face cp integrates to the forces of test_wall_forces_linear_pressure."""
    m = load_example("naca0012_example")
    M = 0.8
    o = Oracle(m, grad_method=3, lsq_nn=1, mach=M, aoa_deg=1.25)
    cl = o.cells()
    pv = np.zeros((o.ncells, 4))
    pv[:, 0] = 1.0
    pv[:, 1] = 0.5 - 0.2 * cl["x"] + 0.1 * cl["y"]
    pv[:, 2] = 0.05 + 0.3 * cl["x"] - 0.4 * cl["y"]
    pv[:, 3] = 2.0 + 0.3 * cl["x"] - 0.7 * cl["y"]
    o.pvar = pv
    faces, un3 = o.wall_cp()
    e, ls = o.edges(), o.lists()
    we = ls["bnd_edges"][: ls["bnd_nedges"][0]]
    assert faces.shape == (we.size, 4)
    xf, yf = e["x"][we], e["y"][we]
    pinf = 1.0 / 1.4
    assert np.allclose(faces[:, 0], xf, rtol=0, atol=0)
    assert np.allclose(faces[:, 1], 2 / M ** 2 * (2.0 + 0.3 * xf - 0.7 * yf - pinf), rtol=1e-12, atol=1e-12)
    un = (0.5 - 0.2 * xf + 0.1 * yf) * e["nx"][we] + (0.05 + 0.3 * xf - 0.4 * yf) * e["ny"][we]
    assert np.allclose(faces[:, 3], un, rtol=1e-12, atol=1e-12)
    bc = m.boundaries[0][1]
    assert np.allclose(faces[:, 2], 2 / M ** 2 * (pv[bc, 3] - pinf), rtol=1e-14, atol=1e-14)
    assert math.isclose(un3[0], np.max(np.abs(un)), rel_tol=1e-12)
    assert math.isclose(un3[1], math.sqrt(np.mean(un ** 2)), rel_tol=1e-12)
    assert math.isclose(un3[2], np.mean(np.abs(un)), rel_tol=1e-12)
    # forces from the face cp in the reference's C_N / C_A sums
    A = e["area"][we]
    cn, ca = np.sum(faces[:, 1] * e["ny"][we] * A), np.sum(faces[:, 1] * e["nx"][we] * A)
    al = math.radians(1.25)
    f = o.wall_forces()
    assert math.isclose(f[0], cn * math.cos(al) - ca * math.sin(al), rel_tol=1e-10)
    assert math.isclose(f[1], cn * math.sin(al) + ca * math.cos(al), rel_tol=1e-10)
