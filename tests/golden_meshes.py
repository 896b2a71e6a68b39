"""Loader of the mesh fixtures in tests/golden/ (see make_fixtures.py)."""
import os

import numpy as np

from fvs2d_amd import meshgen

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_example(name: str) -> meshgen.Mesh:
    z = np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False)
    bnd, off = [], 0
    for n, ln in zip(z["bnames"], z["blens"]):
        bnd.append((str(n), z["bcells"][off:off + ln]))
        off += ln
    return meshgen.Mesh(z["xy"], int(z["ntri"]), int(z["nquad"]), z["cell_nodes"], bnd)
