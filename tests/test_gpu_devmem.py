"""The process-wide cache of large device blocks (devmem.cpp): a context set up on
blocks released by an earlier context (cleared on reuse) gives bit-identical
coefficients and values to the first one, for both precisions."""
import numpy as np
import pytest

import lagrbf_amd as L
from tests import cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_reused_blocks_give_identical_results(precision):
    # 64,000 nodes on a 16^3 PU: ~150-member patches, ~0.75 GB of FP64 factors
    opts, X, V, P = cases.fourier_3d(nlat=40, Np=20000, n=16)
    opts.update(lagrbf_precision=precision, advection=False, lagrbf_setup_all=True)
    print(f"case: device block reuse, {precision}", flush=True)
    out, coefs = [], []
    for k in range(2):
        g = L.LagRBFCtx(petsc_options=dict(opts))
        g.fill_data(X, V); g.fill_points(P); g.interpolate_c()
        v = np.zeros((P.shape[0], 3)); g.output(fvals=v)
        out.append(v)
        coefs.append(np.concatenate([g.debug_get("patch_coeffs", p, dtype=np.float64) for p in range(0, 16 ** 3, 97)]))
        print(f"context {k}: cached before destroy {g.info('device_cache_bytes') / 2**20:.0f} MiB, "
              f"pool growth {g.info('setup_alloc_ms'):.1f} ms", flush=True)
        g.destroy()
        if k == 0:
            g0 = L.LagRBFCtx(petsc_options=dict(opts))
            cached = g0.info("device_cache_bytes")
            g0.destroy()
            print(f"cached after the first context: {cached / 2**20:.0f} MiB")
            assert cached >= 64 * 2**20
    assert np.array_equal(out[0], out[1])
    assert np.array_equal(coefs[0], coefs[1])
