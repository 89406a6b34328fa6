import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def golden():
    def load(name):
        return np.load(os.path.join(GOLDEN, name), allow_pickle=False)
    return load


def seeded_flm(B, seed, red=False):
    """Seeded real-signal flm: f_l0 ~ N(0,1) real, Re/Im f_lm ~ N(0,1) for
    1 <= m <= l, zero for m > l (m < 0 implied by f_{l,-m} = (-1)^m conj f_lm);
    red=True scales by (l+1)^-1."""
    rng = np.random.default_rng(seed)
    f = np.zeros((B, B), dtype=np.complex128, order="F")
    for l in range(B):
        f[l, 0] = rng.standard_normal()
        if l:
            f[l, 1:l + 1] = rng.standard_normal(l) + 1j * rng.standard_normal(l)
        if red:
            f[l, :] /= (l + 1)
    return f
