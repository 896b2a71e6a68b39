"""Brute-force restatements from raw mesh data (test infrastructure).

Everything here starts from a mesh's node coordinates and cell node lists
alone (no topology, stencil or coefficient array of the oracle or of the
product) and computes the published methods in their textbook form, so the
oracle's folded coefficient forms can be checked against them:

* Green-Gauss, cell based: face value by linear interpolation between the two
  cell centres at the face midpoint's distances, d1/(d0+d1) q_i + d0/(d0+d1)
  q_nb (src/gradient_ggcb.f90:84-105), q_i on boundary faces;
* Green-Gauss, node based: node values by inverse-distance weighting of the
  incident cells (src/gradient_ggnb.f90:59-76), face value the mean of its two
  node values, then the Green-Gauss sum -- two passes, where the reference
  folds them into per-cell coefficients (src/gradient_ggnb.f90:82-174);
* the limiter assembly of src/gradient_limiter.f90:47-91 over the
  node-neighbour stencil (cells sharing a node);
* the Roe flux through a slip wall, where the mirrored ghost state makes the
  Roe average closed-form.

Only non-periodic meshes (the oracle's periodic frames are not restated).
Cell centres are vertex averages (src/grid_procs.f90:185-208).
"""
from __future__ import annotations

import math
from collections import defaultdict

import numpy as np


class RawMesh:
    """Cells, centres, volumes and faces of a meshgen.Mesh from its node lists."""

    def __init__(self, mesh):
        assert mesh.canon is None, "brute-force forms cover non-periodic meshes"
        self.xy = np.asarray(mesh.xy, np.float64)
        cells = []
        k = 0
        for _ in range(mesh.ntri):
            cells.append([int(v) for v in mesh.cell_nodes[k:k + 3]])
            k += 3
        for _ in range(mesh.nquad):
            cells.append([int(v) for v in mesh.cell_nodes[k:k + 4]])
            k += 4
        self.cells = cells
        self.N = len(cells)
        self.centre = np.array([self.xy[c].mean(0) for c in cells])
        self.vol = np.array([0.5 * sum(self.xy[c[i]][0] * self.xy[c[(i + 1) % len(c)]][1] -
                                       self.xy[c[(i + 1) % len(c)]][0] * self.xy[c[i]][1] for i in range(len(c)))
                             for c in cells])
        owners = defaultdict(list)
        for ic, c in enumerate(cells):
            for i in range(len(c)):
                owners[frozenset((c[i], c[(i + 1) % len(c)]))].append(ic)
        # per cell: (node p, node q, outward unit normal, length, midpoint, neighbour or -1)
        self.faces = []
        for ic, c in enumerate(cells):
            fl = []
            for i in range(len(c)):
                p, q = c[i], c[(i + 1) % len(c)]
                d = self.xy[q] - self.xy[p]
                L = math.hypot(d[0], d[1])
                n = np.array([d[1], -d[0]]) / L  # counter-clockwise cell: outward
                mid = 0.5 * (self.xy[p] + self.xy[q])
                nb = [j for j in owners[frozenset((p, q))] if j != ic]
                fl.append((p, q, n, L, mid, nb[0] if nb else -1))
            self.faces.append(fl)
        self.node_cells = defaultdict(list)
        for ic, c in enumerate(cells):
            for v in c:
                self.node_cells[v].append(ic)

    def node_stencil(self, ic):
        """Cells sharing a node with ic, ic excluded."""
        s = set()
        for v in self.cells[ic]:
            s.update(self.node_cells[v])
        s.discard(ic)
        return sorted(s)


def ggcb(m: RawMesh, f):
    g = np.zeros((2, m.N))
    for ic in range(m.N):
        acc = np.zeros(2)
        for p, q, n, L, mid, nb in m.faces[ic]:
            if nb < 0:
                qf = f[ic]
            else:
                d0 = np.linalg.norm(mid - m.centre[ic])
                d1 = np.linalg.norm(mid - m.centre[nb])
                qf = (d1 * f[ic] + d0 * f[nb]) / (d0 + d1)
            acc += qf * n * L
        g[:, ic] = acc / m.vol[ic]
    return g


def idw_node_values(m: RawMesh, f, power: float = 1.0):
    out = {}
    for v, cs in m.node_cells.items():
        w = np.array([1.0 / np.linalg.norm(m.centre[c] - m.xy[v]) ** power for c in cs])
        out[v] = float(np.dot(w, f[cs]) / w.sum())
    return out


def ggnb(m: RawMesh, f):
    fv = idw_node_values(m, f)
    g = np.zeros((2, m.N))
    for ic in range(m.N):
        acc = np.zeros(2)
        for p, q, n, L, mid, nb in m.faces[ic]:
            acc += 0.5 * (fv[p] + fv[q]) * n * L
        g[:, ic] = acc / m.vol[ic]
    return g


def venkatakrishnan(a, b, vol, K=5.0):
    """Venkatakrishnan (1993): phi(D+, D-) with eps^2 = (K h)^3, h = 2 sqrt(V/pi)."""
    e2 = (K * 2.0 * math.sqrt(vol / math.pi)) ** 3
    return ((a * a + e2) * b + 2 * b * b * a) / (a * a + 2 * b * b + a * b + e2) / b


def barth(a, b, vol):
    return min(1.0, a / b)


def limiter(m: RawMesh, pvar, grad, fn):
    """phi per cell: for each variable the bounds over the cell and its node
    neighbours, each face midpoint's unlimited increment limited against them,
    the minimum over faces (capped at 1) and then over variables."""
    phi = np.ones(m.N)
    for ic in range(m.N):
        st = m.node_stencil(ic)
        pv = []
        for k in range(4):
            q = pvar[ic, k]
            qs = np.append(pvar[st, k], q)
            dmax, dmin = qs.max() - q, qs.min() - q
            best = 1.0
            for p, qq, n, L, mid, nb in m.faces[ic]:
                diff = (mid[0] - m.centre[ic, 0]) * grad[0, ic, k] + (mid[1] - m.centre[ic, 1]) * grad[1, ic, k]
                if diff > 0:
                    best = min(best, fn(dmax, diff, m.vol[ic]))
                elif diff < 0:
                    best = min(best, fn(dmin, diff, m.vol[ic]))
            pv.append(best)
        phi[ic] = min(pv)
    return phi


def slip_wall_flux(p, nx, ny, g=1.4, delta=0.2):
    """Roe flux between a primitive state p and its mirror image across a wall
    with unit normal n (normal velocity reversed).  The Roe average has zero
    normal velocity, the enthalpy of either side and sound speed
    a_hat^2 = a^2 + (g-1) u_n^2 / 2; the contact and shear waves carry no jump,
    and the two acoustic waves carry equal and opposite strengths, so mass and
    energy fluxes vanish and the momentum flux is p* n with
    p* = p + rho u_n^2 + rho |lambda| u_n, |lambda| = a_hat (Harten's
    (lambda^2/delta + delta)/2 below delta).  Returns (flux, ws_max)."""
    r, u, v, P = p
    un = u * nx + v * ny
    ahat = math.sqrt(g * P / r + 0.5 * (g - 1) * un * un)
    lam = ahat if ahat >= delta else 0.5 * (ahat * ahat / delta + delta)
    pstar = P + r * un * un + r * lam * un
    return np.array([0.0, pstar * nx, pstar * ny, 0.0]), 0.5 * ahat


def phys_flux(p, nx, ny, g=1.4):
    r, u, v, P = p
    un = u * nx + v * ny
    E = P / (g - 1) + 0.5 * r * (u * u + v * v)
    return np.array([r * un, r * un * u + P * nx, r * un * v + P * ny, un * (E + P)])
