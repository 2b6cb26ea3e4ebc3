"""The oracle against its own frozen outputs (tests/golden/oracle_*.npz, written by
tests/golden/make_oracle_fixtures.py from the reference's example files), and the product's
host tables against the frozen LUT slices.  CPU only."""
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN

sys.path.insert(0, GOLDEN)


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


@pytest.fixture(scope="module")
def rebuilt():
    import make_oracle_fixtures
    return make_oracle_fixtures.build()


def test_oracle_reproduces_its_frozen_outputs(rebuilt):
    lut, maps, ev = load("oracle_lut_slices.npz"), load("oracle_maps_1u1b.npz"), load("oracle_eval_1u1b.npz")
    for k in ("smooth_e", "smooth_dor", "fast"):
        assert np.array_equal(rebuilt["lut"][k], lut[k])
    assert np.array_equal(rebuilt["maps"]["n"], maps["n"]) and np.allclose(rebuilt["maps"]["maps"], maps["maps"], rtol=1e-12, atol=1e-12)
    assert np.array_equal(rebuilt["eval"]["conf"], ev["conf"])
    assert np.allclose(rebuilt["eval"]["e"], ev["e"], rtol=1e-12, atol=1e-12)
    assert np.allclose(rebuilt["eval"]["grad"], ev["grad"], rtol=1e-10, atol=1e-10)


def test_frozen_outputs_are_sane():
    """What the frozen numbers must satisfy on their own: fast = mean of adjacent smooth values,
    small tails at the 8 A cutoff; maps finite and attractive somewhere; a spread of energies
    over the poses, with the reference's nine docked modes the lowest (the maps give about -15
    there, as the reference's own scores before the final rescoring do)."""
    lut = load("oracle_lut_slices.npz")
    se, fa = lut["smooth_e"], lut["fast"]
    assert np.allclose(fa[:, :-1], (se[:, :-1] + se[:, 1:]) / 2, rtol=0, atol=1e-15)
    assert np.all(np.abs(se[:, -1]) < 1e-2)
    m = load("oracle_maps_1u1b.npz")["maps"]
    assert m.shape == (17, 9, 9, 9) and np.isfinite(m).all() and m.min() < 0
    ev = load("oracle_eval_1u1b.npz")
    T = ev["conf"].shape[1] - 7
    assert ev["conf"].shape[0] == 64 and ev["grad"].shape == (64, 6 + T)
    e = ev["e"]
    assert np.isfinite(e).all() and e.min() < 0 < e.max()
    assert np.all(e[:9] < -14) and e[:9].max() < e[9:].min()
    assert np.abs(ev["grad"][:9]).max() < 20


def test_product_tables_match_frozen_lut(host_ctx):
    lut = load("oracle_lut_slices.npz")
    for k, tp in enumerate(lut["tp"]):
        se, sd, fa = host_ctx.lut(int(tp))
        assert np.allclose(se, lut["smooth_e"][k], atol=1e-11)
        assert np.allclose(sd, lut["smooth_dor"][k], atol=1e-11)
        assert np.allclose(fa, lut["fast"][k], atol=1e-11)
