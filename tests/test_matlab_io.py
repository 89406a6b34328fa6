"""MATLAB .m/.dat wavelet files (s2dw_amd/fortran/s2dw_fileio_mod.f90)
against files the reference's own writer produced.

tests/golden/ref_mat_* were written by the reference's
s2dw_fileio_matlab_wav_write (src/mod/s2dw_fileio_mod.f90:102-266 static,
298-460 dynamic) through s2dw_hip_fits_test, from tests/golden with relative
names (tests/golden/gen_ref_fixtures.py).  The module here must write the
same three files byte for byte and read the reference's back (readers:
473-615 static, 649-763 dynamic).  No GPU work is involved.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT

FDIR = os.path.join(ROOT, "s2dw_amd", "fortran")
DRIVER = os.path.join(FDIR, "_build", "s2dw_hip_fits_test")
REF = os.path.join(ROOT, "oracle", "_ref", "ref_fits")
GOLDEN = os.path.join(ROOT, "tests", "golden")
CASES = [("dynamic", 8, 2, 2.0, 1), ("static", 6, 2, 3.0, 1)]
PARTS = (".m", "_scoeff.dat", "_wcoeff.dat")

pytestmark = pytest.mark.skipif(shutil.which("amdflang") is None, reason="no Fortran compiler")


def base(layout, B, N, alpha, J):
    return f"ref_mat_{layout}_B{B}_N{N}_a{int(alpha)}_J{J}"


@pytest.fixture(scope="module", autouse=True)
def built():
    r = subprocess.run(["make", "-C", FDIR], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert os.path.exists(DRIVER)


def run(*args, cwd=None):
    cmd = [str(a) for a in args]
    print(" ".join(cmd), flush=True)
    return subprocess.run(cmd, capture_output=True, text=True, timeout=120, cwd=cwd)


def parse_read(stdout):
    head = re.search(r"^header\s+(.*)$", stdout, re.M).group(1).split()
    scales = [line.split() for line in stdout.splitlines() if line.startswith("scale")]
    sc = float(re.search(r"scoeff\s+([0-9.Ee+-]+)", stdout).group(1))
    return head, scales, sc


def band_his(B, alpha, J):
    import oracle as O

    his, _ = O.band_limits(J, B, alpha)
    return his


@pytest.mark.parametrize("layout,B,N,alpha,J", CASES)
def test_reference_files_layout(layout, B, N, alpha, J):
    # the header lines the readers parse (19 skipped, then J, N, B, bl_scoeff,
    # alpha, two skipped, three sizes per scale) and the .dat contents
    name = os.path.join(GOLDEN, base(layout, B, N, alpha, J))
    lines = open(name + ".m").read().splitlines()
    val = lambda s: s.split("=", 1)[1].strip().rstrip(";")
    assert int(val(lines[19])) == J and int(val(lines[20])) == N and int(val(lines[21])) == B
    bl = int(val(lines[22]))
    assert float(val(lines[23])) == alpha
    his = band_his(B, alpha, J)
    sizes = [int(val(lines[26 + k])) for k in range(3 * (J + 1))]
    assert sizes == [s for h in his for s in (2 * h - 1, 2 * h, N)]
    sc = np.loadtxt(name + "_scoeff.dat")
    tri = [(l, m) for l in range(bl) for m in range(l + 1)]
    assert sc.shape == (2 * len(tri),)
    np.testing.assert_array_equal(sc[:len(tri)], [l + 0.5 * m for l, m in tri])
    np.testing.assert_array_equal(sc[len(tri):], [-0.25 * l for l, m in tri])
    w = np.loadtxt(name + "_wcoeff.dat")
    off = 0
    for jj, h in enumerate(his):
        n = (2 * h - 1) * 2 * h * N
        g, b, a = np.meshgrid(np.arange(N), np.arange(2 * h), np.arange(2 * h - 1), indexing="ij")
        np.testing.assert_allclose(w[off:off + n].reshape(N, 2 * h, 2 * h - 1),
                                   np.sin(0.1 * a + 0.2 * b + 0.3 * g + jj), rtol=0, atol=1e-19)
        off += n
    assert off == w.size


@pytest.mark.parametrize("layout,B,N,alpha,J", CASES)
@pytest.mark.parametrize("ours", ["static", "dynamic"])
def test_writer_matches_reference_files(tmp_path, layout, B, N, alpha, J, ours):
    name = base(layout, B, N, alpha, J)
    r = run(DRIVER, "mwrite_" + ours, name + ".m", B, N, alpha, J, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    for part in PARTS:
        assert (tmp_path / (name + part)).read_bytes() == open(os.path.join(GOLDEN, name + part), "rb").read(), part


@pytest.mark.parametrize("layout,B,N,alpha,J", CASES)
@pytest.mark.parametrize("ours", ["static", "dynamic"])
def test_reader_reads_reference_files(layout, B, N, alpha, J, ours):
    r = run(DRIVER, "mread_" + ours, os.path.join(GOLDEN, base(layout, B, N, alpha, J) + ".m"))
    assert r.returncode == 0, r.stdout + r.stderr
    head, scales, sc = parse_read(r.stdout)
    assert [int(v) for v in head[:3]] == [J, B, N] and float(head[4]) == alpha
    his = band_his(B, alpha, J)
    assert len(scales) == J + 1
    for jj, (row, h) in enumerate(zip(scales, his)):
        assert [int(v) for v in row[1:5]] == [jj, 2 * h - 1, 2 * h, N]
        assert float(row[5]) == 0.0
    assert sc == 0.0  # the upper triangle of scoeff reads as zero, as in the reference


@pytest.mark.skipif(not os.path.exists(REF), reason="reference file-I/O module not built (needs reference sources)")
@pytest.mark.parametrize("ours", ["static", "dynamic"])
def test_reference_reader_reads_our_files(tmp_path, ours):
    assert run(DRIVER, "mwrite_" + ours, "w.m", 16, 3, 2.0, 2, cwd=tmp_path).returncode == 0
    for reader in ("mread_static", "mread_dynamic"):
        r = run(REF, reader, tmp_path / "w.m")
        assert r.returncode == 0, r.stdout + r.stderr
        _, scales, sc = parse_read(r.stdout)
        assert all(float(row[5]) == 0.0 for row in scales) and sc == 0.0


def test_existing_file_refused(tmp_path):
    # the reference opens all three files with status='new': an existing one
    # is a Fortran I/O error that ends the program
    assert run(DRIVER, "mwrite_dynamic", "w.m", 8, 2, 2.0, cwd=tmp_path).returncode == 0
    before = (tmp_path / "w.m").read_bytes()
    r = run(DRIVER, "mwrite_static", "w.m", 8, 2, 2.0, cwd=tmp_path)
    assert r.returncode != 0
    assert (tmp_path / "w.m").read_bytes() == before
