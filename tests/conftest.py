import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def c1_disc():
    """The small disc workload: spacing 0.04, seed 0 (~4.5k triangles)."""
    from morphoshell_amd.meshing import disc_mesh

    return disc_mesh(0.04, seed=0)


def grf_disc(spacing=0.04, seed=0, h=0.01, kappa=0.0, amp=0.1, corr=0.2):
    from morphoshell_amd.meshing import disc_mesh, grf, isotropic_abar_info, uniform_bbar_info

    m = disc_mesh(spacing, seed=seed)
    c = m.centroids()
    lam = 1.0 + amp * grf(c[:, :2], corr, seed=seed + 1)
    m.abar_info = isotropic_abar_info(lam)
    m.bbar_info = uniform_bbar_info(m.num_tris, kappa)
    m.fill_defaults(h=h)
    return m


def deformed_state(sh, seed=0, amp=0.05, stretch=1.02):
    """A smooth non-trivial 3-D state near the flat reference."""
    rng = np.random.default_rng(seed)
    N = sh.N
    x = sh.ref_pos.copy()
    X, Y = x[:N].copy(), x[N:2 * N].copy()
    x[:N] = stretch * X
    x[N:2 * N] = Y / stretch
    x[2 * N:] = amp * np.sin(2.5 * X) * np.cos(1.7 * Y) + 0.3 * amp * X * Y + 1e-4 * rng.standard_normal(N)
    v = 1e-3 * rng.standard_normal(3 * N)
    return x, v
