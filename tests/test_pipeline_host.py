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
