"""libs2dw_hip.so on the CPU: it loads, exports every entry point that
include/s2dw_hip.h declares, and its host setup routines (which never touch
the GPU) reproduce the reference's golden vectors.  No GPU compute here."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from s2dw_amd import core

HEADER = os.path.join(ROOT, "include", "s2dw_hip.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(s2dw_hip_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    syms = declared_symbols()
    assert len(syms) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", core.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (\w+)", out))
    missing = [s for s in syms if s not in exported]
    assert not missing, missing
    lib = core.lib()
    for s in syms:
        assert hasattr(lib, s)


def test_version_string():
    assert b"gfx950" in core.lib().s2dw_hip_version()


def test_host_kernels_match_reference(golden):
    g = golden("ref_kernels.npz")
    keys = sorted({k.rsplit("_", 1)[0] for k in g.files if k.endswith("_meta")})
    for key in keys:
        B, J, jmax, bl, ok = (int(x) for x in g[key + "_meta"])
        alpha = float(g[key + "_alpha"][0])
        K, Phi2, bl2 = core.s2dw_core_init_kernels(J, B, alpha)
        assert bl2 == bl
        np.testing.assert_array_max_ulp(K, g[key + "_K"], maxulp=2)
        np.testing.assert_array_max_ulp(Phi2, g[key + "_Phi2"], maxulp=2)
        a, ok2 = core.s2dw_core_admiss(K, Phi2, B, J)
        assert ok2 == bool(ok)
        np.testing.assert_allclose(a, g[key + "_admiss"], rtol=0, atol=1e-15)
        assert core.s2dw_core_comp_Jmax(B, alpha) == jmax
        hi, lo = core.band_limits(J, B, alpha)
        np.testing.assert_array_equal(np.stack([hi, lo], 1), g[key + "_bands"])


def test_host_jmax_and_assimilate(golden):
    for B, alpha, jm in golden("ref_jmax.npy"):
        assert core.s2dw_core_comp_Jmax(int(B), float(alpha)) == int(jm)
    assert core.s2dw_core_assimilate_int(3.000001) == (True, 3)
    assert core.s2dw_core_assimilate_int(2.999995) == (True, 3)
    assert core.s2dw_core_assimilate_int(2.5) == (False, 0)


def test_host_directionality_matches_reference(golden):
    g = golden("ref_direc.npz")
    for key in g.files:
        B, N = (int(x[1:]) for x in key.split("_"))
        np.testing.assert_array_equal(core.s2dw_core_init_directionality(B, N), g[key])


def test_error_codes_mirror_reference():
    # N > B halts in s2dw_core_init_directionality (s2dw_core_mod.f90:228-230) with SIZE_INVALID
    with pytest.raises(core.S2dwError) as e:
        core.s2dw_core_init_directionality(4, 5)
    assert e.value.code == core.S2DW_ERROR_SIZE_INVALID
    assert "Invalid sizes" in str(e.value)
    # J outside [1, J_max] halts in s2dw_core_init_kernels (112-114)
    with pytest.raises(core.S2dwError) as e:
        core.s2dw_core_init_kernels(9, 64, 2.0)
    assert e.value.code == core.S2DW_ERROR_SIZE_INVALID
    with pytest.raises(core.S2dwError):
        core.s2dw_core_init_kernels(0, 64, 2.0)


def test_admissibility_failure_detected():
    K, Phi2, _ = core.s2dw_core_init_kernels(6, 64, 2.0)
    K2 = K.copy()
    K2[2, 20] *= 1.001
    _, ok = core.s2dw_core_admiss(K2, Phi2, 64, 6)
    assert not ok
