"""DELTA_TIMING stage timers (src/timing.cpp:19-46, src/bindings.cpp:591-599):
with the variable set, each call's result dict carries {label: seconds} of
the device stages it ran, under the reference's DELTA_TIME labels where the
stage has one (e.g. "subtract: model convolve"), and drain_timing() returns
them for calls that return bare arrays; with it unset or "0" the dicts carry
None."""
import numpy as np
import pytest

from tests.test_gpu_fit import _pair

pytestmark = pytest.mark.gpu


def _run(core):
    p, sx, sy = _pair()
    n = p["science"].shape[0]
    knots = np.array([[x, y] for y in (0.0, n - 1.0) for x in (0.0, n - 1.0)] +
                     [[0.5 * (n - 1), 0.5 * (n - 1)]])
    fit = core.fit_kernel(p["science"], p["reference"], knots, sx, sy, 10, 1.0, 3,
                          np.logspace(-4, 4, 9), radius=10, science_var=p["science_var"],
                          reference_var=p["reference_var"])
    sub = core.subtract(p["science"], p["reference"], knots, fit["theta"], 1.0, 3, radius=10,
                        saturation=6e4, science_var=p["science_var"],
                        reference_var=p["reference_var"])
    dec = core.decorrelate(sub["difference"], knots, fit["theta"], 1.0, 3, p["science_var"],
                           p["reference_var"], block=64, radius=10)
    psf = np.ones((5, 5), np.float32) / 25
    core.matched_filter(dec["difference"], psf, dec["variance"])
    return fit, sub, dec, core.drain_timing()


def test_timing_dicts_under_delta_timing(gpu, monkeypatch):
    from delta_amd import _core

    monkeypatch.setenv("DELTA_TIMING", "1")
    monkeypatch.setattr(_core, "_TIMING_ON", None)
    print("DELTA_TIMING=1", flush=True)
    fit, sub, dec, mf = _run(_core)
    for d, labels in ((fit["timing"], ["fit: stamp B_n convolve", "fit: normal equations",
                                       "fit: lambda grid solves"]),
                      (sub["timing"], ["subtract: sanitise reference", "subtract: spatial fields",
                                       "subtract: model convolve",
                                       "subtract: variance propagation",
                                       "subtract: mask growth"]),
                      (dec["timing"], ["decorrelate: kernel-power cache",
                                       "decorrelate: block filter+blend",
                                       "decorrelate: normalise"]),
                      (mf, ["score: blocked correlation"])):
        print(d, flush=True)
        assert isinstance(d, dict)
        for k in labels:
            assert d.get(k, 0.0) > 0.0, k
        assert all(isinstance(v, float) and v >= 0.0 for v in d.values())
    # a call clears the previous one's timers
    assert "fit: normal equations" not in sub["timing"]
    assert "subtract: model convolve" not in mf


def test_timing_off(gpu, monkeypatch):
    from delta_amd import _core

    monkeypatch.setenv("DELTA_TIMING", "0")
    monkeypatch.setattr(_core, "_TIMING_ON", None)
    fit, sub, dec, mf = _run(_core)
    assert fit["timing"] is None and sub["timing"] is None and dec["timing"] is None
    assert mf is None
