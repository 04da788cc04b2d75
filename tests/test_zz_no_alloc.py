"""The allocation-free realft (rpmd_realft_noalloc) against the context's rpmd_realft, through the
C-ABI: two calls in a row, the second with a smaller input, both directions, one of them over
several row chunks of the fixed device array; and the error path for n beyond the maximum."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_realft_noalloc_matches_realft():
    from rpmd_mainz_amd import capi
    from rpmd_mainz_amd import configs as cf
    rng = np.random.default_rng(17)
    cases = [(3 * 21000, 32, +1), (999, 7, -1)]     # 63000 rows: four chunks of 16384 at n = 32
    prm = cf.make_params(cf.MODELS["qspcfw"], [20.0, 20.0, 20.0])
    ctx = capi.RPMDContext(prm, 21000, 32, "fp64")
    for m, n, mode in cases:
        print(f"case noalloc realft m {m} n {n} mode {mode}", flush=True)
        x = rng.normal(size=(n, m))
        a = capi.realft_noalloc(x.copy(), m, n, mode)
        b = ctx.realft(x.copy(), m, n, mode)
        scale = np.abs(b).max()
        assert np.abs(a - b).max() <= 1e-15 * scale, (m, n, mode)
        # and the transform itself: R2HC (mode +1) of x, or its inverse, scaled by 1/sqrt(n)
        X = np.fft.fft(x, axis=0) / np.sqrt(n) if mode == 1 else None
        if X is not None:
            assert np.allclose(a[: n // 2 + 1], X[: n // 2 + 1].real, atol=1e-12 * scale)
    ctx.close()
    with pytest.raises(capi.RPMDError):
        capi.realft_noalloc(np.zeros((257, 3)), 3, 257, 1)
