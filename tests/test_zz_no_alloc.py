"""GPU: the allocation-free labelling entry point (dg_label_device_noalloc)
against dg_label_device, called through the C ABI: two calls in a row, the
second on a smaller frame (the static scratch keeps the first call's
contents), both polarities, a mask, and the size limit."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _lib():
    from delta_amd import _core

    lib = C.CDLL(_core.library_path())
    lib.dg_label_device_noalloc.restype = C.c_int
    lib.dg_label_device_noalloc.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                            C.c_double, C.c_int, C.c_void_p,
                                            C.POINTER(C.c_int), C.c_void_p]
    lib.dg_last_error.restype = C.c_char_p
    return lib


def _score(h, w, seed):
    rng = np.random.default_rng(seed)
    s = rng.standard_normal((h, w)).astype(np.float32)
    for _ in range(h * w // 3000):
        y, x = rng.integers(0, h), rng.integers(0, w)
        r = int(rng.integers(1, 6))
        s[max(0, y - r):y + r + 1, max(0, x - r):x + r + 1] += rng.choice([-9.0, 9.0])
    s[rng.integers(0, h, 20), rng.integers(0, w, 20)] = np.nan
    return s


def _noalloc(lib, score, mask, thr, pol, stream):
    import torch

    h, w = score.shape
    out = torch.full((h, w), -7, dtype=torch.int32, device=score.device)
    n = C.c_int(-1)
    rc = lib.dg_label_device_noalloc(w, h, score.data_ptr(),
                                     None if mask is None else mask.data_ptr(), thr, pol,
                                     out.data_ptr(), C.byref(n), stream.cuda_stream)
    assert rc == 0, lib.dg_last_error()
    return out, n.value


def test_label_noalloc_matches_label_device(gpu):
    import torch

    from delta_amd import _core

    lib = _lib()
    stream = torch.cuda.current_stream(gpu)
    cases = [(2100, 1900, 1, False, 1), (300, 257, 2, True, -1), (1, 1, 3, False, 1)]
    for h, w, seed, masked, pol in cases:
        print(f"label_noalloc {h}x{w} polarity {pol} mask {masked}", flush=True)
        score = torch.from_numpy(_score(h, w, seed)).to(gpu)
        mask = None
        if masked:
            m = (np.random.default_rng(seed).random((h, w)) < 0.05).astype(np.uint8)
            mask = torch.from_numpy(m).to(gpu)
        want, n_want = _core.label_components_device(score, 5.0, pol, mask=mask)
        got, n_got = _noalloc(lib, score, mask, 5.0, pol, stream)
        torch.cuda.synchronize()
        assert n_got == n_want
        assert torch.equal(got, want)
    # a frame over 2^28 pixels is refused before anything runs
    big = torch.zeros((16384, 16385), dtype=torch.float32, device=gpu)
    lab = torch.zeros((16384, 16385), dtype=torch.int32, device=gpu)
    n = C.c_int(0)
    rc = lib.dg_label_device_noalloc(16385, 16384, big.data_ptr(), None, 5.0, 1, lab.data_ptr(),
                                     C.byref(n), stream.cuda_stream)
    assert rc != 0 and b"2^28" in lib.dg_last_error()
