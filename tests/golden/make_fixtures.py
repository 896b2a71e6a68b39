"""Write the reference's example meshes and input decks as test fixtures.

Reads (in the build container only) the data files the reference ships under
examples/ and stores them here: meshes as compressed .npz (0-based ids), input
decks as text copies.  Run: python tests/golden/make_fixtures.py
"""
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from fvs2d_amd import meshgen  # noqa: E402

REF = "/root/reference/examples"
CASES = {"vortex_example": ("isentropic_vortex", "vortex"),
         "naca0012_example": ("naca0012_ogrid", "naca0012_omesh")}

for name, (d, base) in CASES.items():
    m = meshgen.read_grid(os.path.join(REF, d, base))
    names = np.array([b[0] for b in m.boundaries])
    lens = np.array([len(b[1]) for b in m.boundaries], np.int32)
    cells = np.concatenate([b[1] for b in m.boundaries]).astype(np.int32)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), xy=m.xy, ntri=m.ntri, nquad=m.nquad,
                        cell_nodes=m.cell_nodes, bnames=names, blens=lens, bcells=cells)
    for f in ("fvs2d.input", "fvs2d.vortex"):
        src = os.path.join(REF, d, f)
        if os.path.exists(src):
            shutil.copy(src, os.path.join(HERE, f"{name}.{f}"))
    print(name, m.nnodes, m.ntri, m.nquad, [(b[0], len(b[1])) for b in m.boundaries])
