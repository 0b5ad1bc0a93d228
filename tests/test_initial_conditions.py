"""Initial-condition generators of the workloads (ow3d/problems.py): the JONSWAP spectrum against the
reference's own JONSWAP_spectrum (utilities/JONSWAP_spectrum.f90:1-30), the linear dispersion relation,
and the significant wave height of the synthesized random surface.

The reference's spectrum takes frequency in Hz and scales with alpha = 3.32285 Hs^2 f_p^4
(JONSWAP_spectrum.f90:13-16); ours is written in angular frequency and rescaled so that 4 sqrt(m0) = Hs
(random_waves), so the two must agree up to one constant factor at every frequency.  Golden values
(tests/golden/jonswap_ref.npz) come from the compiled reference routine (tests/golden/make_golden.py).
"""
import math
import os

import numpy as np
import pytest

import oracle

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _shape_ratio(f, S_ref, tp, gam):
    from ow3d import problems
    m = S_ref > 1e-250                 # far below the peak both underflow to zero
    mine = problems.jonswap(2 * np.pi * f[m], 2 * np.pi / tp, gam)
    r = mine / S_ref[m]
    return np.max(np.abs(r / np.median(r) - 1.0)), int(m.sum())


def test_jonswap_shape_matches_the_reference_golden_values():
    with np.load(os.path.join(HERE, "jonswap_ref.npz"), allow_pickle=False) as z:
        for q in range(2):
            hs, tp, gam = z[f"par{q}"]
            dev, used = _shape_ratio(z[f"f{q}"], z[f"S{q}"], tp, gam)
            assert used > 300 and dev < 1e-11, (q, dev)


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref (compiled reference) not built")
def test_jonswap_shape_matches_the_compiled_reference():
    rng = np.random.default_rng(4)
    for tp, gam in ((1.0, 3.3), (7.5, 2.0), (3.0, 5.0)):
        f = np.sort(rng.uniform(0.3, 5.0, 200)) / tp
        dev, _ = _shape_ratio(f, oracle.ref_jonswap(f, 0.1, tp, gam), tp, gam)
        assert dev < 1e-11, (tp, gam, dev)


def test_wavenumber_solves_the_linear_dispersion_relation():
    from ow3d import problems
    for h in (0.5, 1.0, 2.0):
        for om in (0.3, 1.0, 3.65, 10.0):
            k = problems.wavenumber(om, h)
            assert abs(problems.G * k * math.tanh(k * h) - om * om) < 1e-12 * om * om


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_synthesized_surface_reaches_the_requested_significant_height(seed):
    """Hs = 4 sqrt(m0): the spatial variance of the 257 x 257 JONSWAP surface (64 components over
    16 peak wavelengths) must give the requested Hs = 0.02 h within 2 %."""
    from ow3d import problems
    cfg, bath, E, P, dt, _ = problems.workload("tank257", seed=seed)
    e = E[1:-1, 1:-1]
    hs = 4.0 * math.sqrt(float(np.mean((e - e.mean()) ** 2)))
    assert abs(hs / 0.02 - 1.0) < 0.02, hs
