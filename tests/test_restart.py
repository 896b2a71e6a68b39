"""Restart files (.cd / .s8 / .s4) against the reference's own module.

oracle/_ref/ios_ref is src/ios_unstrc.f90 compiled from the reference's
sources under a test driver (oracle/ios_driver.f90) that mirrors the
reference's save (src/io.f90:96-113, 166-178) and restart
(src/initialize.f90:56-77) call sequences.  The files this package writes
must be byte-identical to the module's, and each side must read the other's
files exactly.  Without the reference sources (or amdflang) only the round
trip and the error paths run.
"""
import os
import subprocess

import numpy as np
import pytest

from fvs2d_amd import restart
from oracle.oracle import IOS_REF

need_ref = pytest.mark.skipif(not os.path.exists(IOS_REF), reason="reference restart module not built")


def _state(n, seed=0):
    rng = np.random.default_rng(seed)
    return np.column_stack([rng.uniform(0.5, 1.5, n), rng.normal(0, 0.5, n), rng.normal(0, 0.5, n),
                            rng.uniform(1.5, 3.0, n)])


def _ref(*args, cwd):
    subprocess.run([IOS_REF, *map(str, args)], cwd=cwd, check=True, capture_output=True)


def test_round_trip_and_header(tmp_path):
    cv = _state(37)
    base = str(tmp_path / "save")
    restart.write_save(base, cv, nnodes=29, itime=1234)
    back = restart.read_cont(base, 37, 29)
    assert np.array_equal(back, cv)
    cd = restart.read_cd(base + ".cd")
    assert (cd["m1"], cd["m2"], cd["nparams"], cd["ntimes"], cd["itimes"]) == (37, 29, 4, 1, [1234])
    assert cd["info_params"] == list(restart.PARAMS)
    raw = np.fromfile(base + ".s8", ">f8")
    assert raw.size == 4 * 37 and raw[0] == cv[0, 0] and raw[37] == cv[0, 1]   # one record per variable


def test_restart_errors(tmp_path):
    base = str(tmp_path / "cont")
    restart.write_save(base, _state(10), nnodes=8, itime=1)
    with pytest.raises(ValueError, match="dimension"):
        restart.read_cont(base, 11, 8)
    with pytest.raises(ValueError, match="dimension"):
        restart.read_cont(base, 10, 9)
    with open(base + ".cd", "w") as f:
        f.write(restart.cd_text(10, 8, [1, 2], ["x"], restart.PARAMS))
    with pytest.raises(ValueError, match="nt>1"):
        restart.read_cont(base, 10, 8)


@need_ref
@pytest.mark.parametrize("imach", [2, 1])
def test_files_identical_to_reference_module(tmp_path, imach):
    n, nn, it = 53, 41, 7
    cv = _state(n, 3)
    cv.T.astype(">f8").tofile(tmp_path / "in.be")
    _ref("write", "refsave", imach, n, nn, it, "in.be", cwd=tmp_path)
    restart.write_save(str(tmp_path / "mine"), cv, nnodes=nn, itime=it, imach=imach)
    ext = restart.EXT[imach]
    assert (tmp_path / "mine.cd").read_bytes() == (tmp_path / "refsave.cd").read_bytes()
    assert (tmp_path / ("mine" + ext)).read_bytes() == (tmp_path / ("refsave" + ext)).read_bytes()


@need_ref
def test_each_side_reads_the_other(tmp_path):
    n, nn = 64, 50
    cv = _state(n, 4)
    # this package writes, the reference reads
    restart.write_save(str(tmp_path / "cont"), cv, nnodes=nn, itime=99)
    _ref("read", "cont", 2, "out.be", cwd=tmp_path)
    out = np.fromfile(tmp_path / "out.be", ">f8")
    assert list(out[:5]) == [n, nn, 4, 1, 99]
    assert np.array_equal(out[5:].reshape(4, n).T, cv)
    # the reference writes, this package reads
    cv2 = _state(n, 5)
    cv2.T.astype(">f8").tofile(tmp_path / "in.be")
    _ref("write", "save", 2, n, nn, 3, "in.be", cwd=tmp_path)
    assert np.array_equal(restart.read_cont(str(tmp_path / "save"), n, nn), cv2)
