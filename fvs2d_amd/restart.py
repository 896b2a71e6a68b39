"""The reference's restart files: ``<base>.cd`` (control text) and
``<base>.s8`` / ``<base>.s4`` (big-endian float64 / float32 records).

Format, as src/ios_unstrc.f90 writes and reads it (writecd 141-298, write8
342-375, readcd 438-570, read8 619-651) with the reference's Intel build
flags (``-convert big_endian``, record length in 4-byte units):

* ``.cd``: fixed-column text -- ``number of nodes = n``, ``number of cells =
  n``, ``number of parameters`` and ``number of timesteps`` (i5), a blank line,
  ``Information about file :   (nnn  info lines )``, the info lines (3x, a72),
  ``Information about parameters :``, the parameter names (3x, a72),
  ``Numbers of timesteps :`` and the step numbers (6 per line, 2x i10);
* data file: direct access, one record per (time level, parameter) in that
  order, each record the m1 values of one parameter.

The save file of a run (src/io.f90:96-113, write_save_ios 166-178) stores the
conservative variables at cell centres, one record each (rho, rhou, rhov,
rhoE), and, as the reference notes, puts ncells in the "number of nodes" line
and nnodes in the "number of cells" line.  A restart (src/initialize.f90:56-77)
reads ``cont.cd`` / ``cont.s8`` back with the same convention and stops when
the sizes differ from the grid's or the file holds more than one time level or
other than four parameters.

These files let the reference's own Fortran driver start from a state this
library produced (a converged batch instance, a warm start).
"""
from __future__ import annotations

import os
from typing import List, Sequence

import numpy as np

EXT = {1: ".s4", 2: ".s8"}
DTYPE = {1: ">f4", 2: ">f8"}
PARAMS = ("rho", "rhou", "rhov", "rhoE")


def _a72(s: str) -> str:
    return "   " + s[:72].ljust(72)


def cd_text(m1: int, m2: int, itimes: Sequence[int], info_file: Sequence[str],
            info_params: Sequence[str]) -> str:
    """Text of a control file (writecd, src/ios_unstrc.f90:236-251): m1 goes
    in the "number of nodes" line, m2 in the "number of cells" line."""
    lines = [f"     number of nodes = {m1}", f"     number of cells = {m2}",
             f"     number of parameters = {len(info_params):5d}", f"     number of timesteps  = {len(itimes):5d}",
             "", f"     Information about file :   ({len(info_file):3d}  info lines )"]
    lines += [_a72(s) for s in info_file]
    lines.append("      Information about parameters :")
    lines += [_a72(s) for s in info_params]
    lines.append("  Numbers of timesteps :")
    its = list(itimes)[:5000]
    for k in range(0, len(its), 6):
        lines.append("".join(f"  {int(t):10d}" for t in its[k:k + 6]))
    return "\n".join(lines) + "\n"


def read_cd(path: str) -> dict:
    """Parse a control file at the columns readcd reads
    (src/ios_unstrc.f90:491-500)."""
    with open(path) as f:
        lines = f.read().split("\n")
    m1 = int(lines[0][23:])
    m2 = int(lines[1][23:])
    mp = int(lines[2][28:33])
    mt = int(lines[3][28:33])
    minf = int(lines[5][33:36])
    k = 6
    info_file = [ln[3:].rstrip() for ln in lines[k:k + minf]]
    k += minf + 1
    info_params = [ln[3:].rstrip() for ln in lines[k:k + mp]]
    k += mp + 1
    vals: List[int] = []
    while len(vals) < min(mt, 5000) and k < len(lines):
        vals += [int(v) for v in lines[k].split()]
        k += 1
    return dict(m1=m1, m2=m2, nparams=mp, ntimes=mt, itimes=vals[:mt], info_file=info_file,
                info_params=info_params)


def write_data(path: str, records, imach: int = 2) -> None:
    """Records (time level major, then parameter) of m1 values each."""
    np.ascontiguousarray(records, np.float64).astype(DTYPE[imach]).tofile(path)


def read_data(path: str, m1: int, nparams: int, ntimes: int = 1, imach: int = 2) -> np.ndarray:
    """(ntimes, nparams, m1) float64 from a data file; record (t, p) starts at
    value ((t-1)*nparams + p-1)*m1 (read8, src/ios_unstrc.f90:625-640)."""
    a = np.fromfile(path, DTYPE[imach], count=ntimes * nparams * m1)
    if a.size != ntimes * nparams * m1:
        raise ValueError(f"{path}: {a.size} values, expected {ntimes * nparams * m1}")
    return a.astype(np.float64).reshape(ntimes, nparams, m1)


def write_save(base: str, cvar, nnodes: int, itime: int, imach: int = 2) -> None:
    """Save file of a run (io_init, src/io.f90:96-113, and write_save_ios,
    166-178): cvar is (ncells, 4); itime the step number recorded
    (ntstart - 1 + ntimes in the reference)."""
    cvar = np.asarray(cvar, np.float64)
    ncells = cvar.shape[0]
    info = [f"number of time-step computed = {itime:6d}", "conservatve variables are saved in cell centers",
            "#ncells and #nodes are replaced", " "]
    with open(base + ".cd", "w") as f:
        f.write(cd_text(ncells, nnodes, [itime], info, PARAMS))
    write_data(base + EXT[imach], cvar.T, imach)


def read_cont(base: str, ncells: int, nnodes: int, imach: int = 2) -> np.ndarray:
    """Restart state (initialize_solution, src/initialize.f90:56-77): (ncells,
    4) conservative variables.  Raises ValueError where the reference stops."""
    cd = read_cd(base + ".cd")
    if cd["m1"] != ncells or cd["m2"] != nnodes:
        raise ValueError("dimension between grid and cont files does not match!")
    if cd["ntimes"] > 1 or cd["nparams"] != 4:
        raise ValueError("nt>1 or #s parameters/=nvar!")
    if not os.path.exists(base + EXT[imach]):
        raise ValueError(f"{base + EXT[imach]} does not exist")
    return read_data(base + EXT[imach], ncells, 4, 1, imach)[0].T.copy()
