"""Host logic of the caller layer (no GPU): input coercion as the reference's
python/delta/_inputs.py:18-72 does it, the knot-count rule of
pipeline.py:504-509, and the configuration/call keyword split of
pipeline.py:911-940."""
import numpy as np
import pytest


def _pl():
    from delta_amd import pipeline
    return pipeline


class _Unc:
    def __init__(self, array, kind):
        self.array = array
        self.uncertainty_type = kind


class _CCD:
    def __init__(self, data, unc=None, mask=None):
        self.data = data
        self.uncertainty = unc
        self.mask = mask


def test_as_layers_plain_and_overrides():
    a = np.arange(12, dtype=np.float64).reshape(3, 4)
    d, v, m = _pl().as_layers(a)
    assert d.dtype == np.float32 and d.flags.c_contiguous and v is None and m is None
    d, v, m = _pl().as_layers(a, variance=np.ones((3, 4)), mask=np.eye(3, 4, dtype=bool))
    assert v.dtype == np.float32 and m.dtype == np.uint8 and m.sum() == 3
    with pytest.raises(ValueError, match="2-D"):
        _pl().as_layers(np.zeros(5))
    with pytest.raises(ValueError, match="variance shape"):
        _pl().as_layers(a, variance=np.ones((4, 3)))
    with pytest.raises(ValueError, match="mask shape"):
        _pl().as_layers(a, mask=np.ones((2, 2), np.uint8))


def test_as_layers_astropy_like():
    data = np.ones((4, 5), np.float32)
    sig = np.full((4, 5), 3.0)
    d, v, m = _pl().as_layers(_CCD(data, _Unc(sig, "std"), np.zeros((4, 5), bool)))
    np.testing.assert_array_equal(v, np.full((4, 5), 9.0, np.float32))
    assert m.dtype == np.uint8 and not m.any()
    d, v, m = _pl().as_layers(_CCD(data, _Unc(sig, "var")))
    np.testing.assert_array_equal(v, sig.astype(np.float32))
    ma = np.ma.MaskedArray(data)  # mask is the scalar nomask
    d, v, m = _pl().as_layers(ma)
    assert m is None


def test_as_layers_masked_array_keeps_its_mask():
    """A MaskedArray is an ndarray subclass; its .data and per-pixel .mask are
    read as _inputs.py:18-38 reads them (True == bad -> 1), and an explicit
    mask argument still overrides it."""
    data = np.arange(20, dtype=np.float64).reshape(4, 5)
    bad = np.zeros((4, 5), bool)
    bad[1, 2] = bad[3, 0] = True
    ma = np.ma.MaskedArray(data, mask=bad)
    d, v, m = _pl().as_layers(ma)
    np.testing.assert_array_equal(d, data.astype(np.float32))
    assert v is None and m.dtype == np.uint8
    np.testing.assert_array_equal(m, bad.astype(np.uint8))
    over = np.zeros((4, 5), np.uint8)
    _, _, m2 = _pl().as_layers(ma, mask=over)
    assert not m2.any()


def test_default_beta_bit_equal_to_reference_order():
    """_default_beta (python/delta/pipeline.py:267-282) scales each FWHM by the
    rounded constant 1/2.354820045030949 before squaring.  Over many FWHM
    pairs the package's beta equals that order bit for bit, while dividing by
    2.354820045030949 instead -- a plausible slip -- differs for some pairs."""
    from delta_amd import synth

    to_sigma = 1.0 / 2.354820045030949
    rng = np.random.default_rng(7)
    pairs = rng.uniform(0.5, 12.0, (20000, 2))
    n_div_differs = 0
    for a, b in pairs:
        a, b = float(a), float(b)
        hi, lo = max(a, b), min(a, b)
        dd = (hi * to_sigma) ** 2 - (lo * to_sigma) ** 2
        want = float(max(0.8 * (np.sqrt(dd) if dd > 0 else 0.0), 0.5))
        assert synth.default_beta(a, b) == want, (a, b)
        dv = (hi / 2.354820045030949) ** 2 - (lo / 2.354820045030949) ** 2
        alt = float(max(0.8 * (np.sqrt(dv) if dv > 0 else 0.0), 0.5))
        n_div_differs += alt != want
    assert n_div_differs > 100
    assert synth.default_beta(float("nan"), -1.0) == 2.0
    assert synth.default_beta(3.0, 3.0) == 0.5          # equal PSFs hit the floor
    assert synth.default_beta(2.5, float("inf")) == 0.5   # one usable FWHM


def test_knot_counts():
    kc = _pl().knot_counts
    assert kc(4096, 4096, 5) == (5, 5)
    assert kc(4096, 2048, 5, 1500.0) == (4, 3)    # round(4095/1500)+1 = 4; 2047/1500 -> 2 -> 3
    assert kc(1000, 1000, 5, 10.0) == (5, 5)       # capped at n_knots
    assert kc(1000, 1000, 5, 0.0) == (5, 5)        # non-positive scale keeps n_knots


def test_config_keys_match_constructor():
    import inspect

    params = set(inspect.signature(_pl().Subtractor.__init__).parameters) - {"self"}
    assert params == set(_pl()._CONFIG_KEYS)
    s = _pl().Subtractor()
    assert s.cv_folds == 5 and not s.decorrelate and not s.score and s.source_catalog
    np.testing.assert_array_equal(s.lambda_grid, np.logspace(-6.0, 6.0, 25))
    cards = s.config_cards()
    assert cards["DLTCVF"] == 5 and cards["DLTSSCL"] == 0.0 and len(cards) == 16


def test_delta_timing_switch(monkeypatch):
    """DELTA_TIMING is on when set, non-empty and not "0" (src/timing.cpp:19-25);
    drain_timing() is None when it is off."""
    from delta_amd import _core

    for value, on in (("1", True), ("yes", True), ("0", False), ("", False), (None, False)):
        if value is None:
            monkeypatch.delenv("DELTA_TIMING", raising=False)
        else:
            monkeypatch.setenv("DELTA_TIMING", value)
        monkeypatch.setattr(_core, "_TIMING_ON", None)
        assert _core.timing_enabled() is on, value
        if not on:
            assert _core.drain_timing() is None
