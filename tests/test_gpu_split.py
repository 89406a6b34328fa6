"""GPU tests of the split-precision form (S2DW_PREC_SPLIT): the analysis
l-contraction on int8 MFMA with exact integer slice products (Ozaki-style
splitting, FP64 recombination), against the FP64 oracle.

Bound: 1e-10 max abs error per scale on the wavelet coefficients and on the
reconstruction (the synthesis of the split analysis's output).  The error of
the form itself is set by its 7 base-256 digits per operand, 2^-54 of the
row's largest |f_lm| per product."""
import sys

import numpy as np
import pytest

import oracle as O
from conftest import seeded_flm
from s2dw_amd import core

pytestmark = pytest.mark.gpu

TOL = 1e-10


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def split_round(B, N, alpha, J, f):
    import torch

    st = O.Setup(B, N, alpha, J=J)
    plan = core.Plan(B, N, st.J, alpha, st.K, st.Phi2, st.S, precision=core.PREC_SPLIT)
    d_f = torch.from_numpy(np.ascontiguousarray(f.T)).cuda()
    d_w, d_sc, d_g = plan.empty_wav(), plan.empty_scoeff(), plan.empty_flm()
    plan.analysis(d_f, d_w, d_sc)
    plan.synthesis(d_w, d_sc, d_g)
    plan.sync()
    w = [np.transpose(plan.scale_view(d_w, j).cpu().numpy(), (2, 1, 0)) for j in range(st.J + 1)]
    sc = d_sc.cpu().numpy().T
    g = d_g.cpu().numpy().T
    plan.close()
    return st, w, sc, g


@pytest.mark.parametrize("B,N,alpha,J,red", [(16, 3, 2.0, None, False), (32, 1, 2.0, None, False),
                                             (64, 5, 2.0, None, False), (27, 4, 3.0, None, True),
                                             (128, 8, 2.0, 5, True), (100, 6, 3.0, None, False)])
def test_split_vs_oracle(B, N, alpha, J, red):
    say(f"split B={B} N={N} alpha={alpha} J={J} red={red}")
    f = seeded_flm(B, 300 + B + N, red=red)
    st, w, sc, g = split_round(B, N, alpha, J, f)
    w_ref, sc_ref = st.analysis(f)
    errs = [float(np.abs(a - b).max()) for a, b in zip(w, w_ref)]
    e_rt = float(np.abs(g - f).max())
    say("  wav err per scale", ["%.1e" % e for e in errs], f"round trip {e_rt:.1e}")
    np.testing.assert_array_equal(sc, sc_ref)
    assert max(errs) <= TOL and e_rt <= TOL


def test_split_L512_N8_Jmin2_per_scale():
    """The benchmark configuration: L=512, N=8, alpha=2, J=7, red spectrum;
    the error against the FP64 oracle is stated per scale."""
    B, N, alpha, J = 512, 8, 2.0, 7
    say("split L=512 N=8 J=7 red (oracle ~10 s)")
    f = seeded_flm(B, 1234, red=True)
    st, w, sc, g = split_round(B, N, alpha, J, f)
    w_ref, sc_ref = st.analysis(f)
    errs = [float(np.abs(a - b).max()) for a, b in zip(w, w_ref)]
    e_rt = float(np.abs(g - f).max())
    say("  wav err per scale", ["%.2e" % e for e in errs], f"round trip {e_rt:.2e}")
    assert max(errs) <= TOL and e_rt <= TOL


def test_split_refused_above_eight_columns():
    B, N = 32, 9
    st = O.Setup(B, N, 2.0)
    with pytest.raises(core.S2dwError) as e:
        core.Plan(B, N, st.J, 2.0, st.K, st.Phi2, st.S, precision=core.PREC_SPLIT)
    assert e.value.code == core.S2DW_ERROR_ARG_INVALID
