"""The process-wide cache of device memory (devmem.cpp): a context set up on physical
chunks released by an earlier context (mapped again and cleared on reuse) gives
bit-identical coefficients and values to the first one, for both precisions."""
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


def test_chunks_reused_by_a_context_of_another_size():
    # a larger context after a smaller one maps the smaller one's chunks and new ones;
    # its values match a second context of the same size set up on its own chunks
    small = cases.fourier_3d(nlat=30, Np=5000, n=12)
    large = cases.fourier_3d(nlat=40, Np=5000, n=16)
    print("case: chunk reuse across context sizes, fp64", flush=True)

    def run(case):
        opts, X, V, P = case
        opts = dict(opts, lagrbf_precision="fp64", advection=False, lagrbf_setup_all=True)
        g = L.LagRBFCtx(petsc_options=opts)
        g.fill_data(X, V); g.fill_points(P); g.interpolate_c()
        v = np.zeros((P.shape[0], 3)); g.output(fvals=v)
        held = g.info("device_cache_bytes")
        g.destroy()
        return v, held

    run(small)
    g0 = L.LagRBFCtx(petsc_options=dict(small[0]))
    before = g0.info("device_cache_bytes")
    g0.destroy()
    v1, held1 = run(large)
    v2, _ = run(large)
    print(f"cached after the small context {before / 2**20:.0f} MiB, while the large one ran "
          f"{held1 / 2**20:.0f} MiB", flush=True)
    assert before >= 64 * 2**20
    assert held1 < before   # the large context took the small one's chunks
    assert np.array_equal(v1, v2)


def test_cache_off_gives_identical_values(tmp_path):
    """LAGRBF_DEVICE_CACHE=0 (plain hipMalloc / hipFree for every pool) in a child
    process gives the same values bit for bit as the cached ranges here."""
    import os
    import subprocess
    import sys
    opts, X, V, P = cases.fourier_3d(nlat=40, Np=20000, n=16)
    opts.update(lagrbf_precision="mixed", advection=False, lagrbf_setup_all=True)
    g = L.LagRBFCtx(petsc_options=dict(opts))
    g.fill_data(X, V); g.fill_points(P); g.interpolate_c()
    here = np.zeros((P.shape[0], 3)); g.output(fvals=here)
    g.destroy()
    out = tmp_path / "off.npy"
    code = (
        "import numpy as np, lagrbf_amd as L\n"
        "from tests import cases\n"
        "opts, X, V, P = cases.fourier_3d(nlat=40, Np=20000, n=16)\n"
        "opts.update(lagrbf_precision='mixed', advection=False, lagrbf_setup_all=True)\n"
        "g = L.LagRBFCtx(petsc_options=dict(opts))\n"
        "g.fill_data(X, V); g.fill_points(P); g.interpolate_c()\n"
        "v = np.zeros((P.shape[0], 3)); g.output(fvals=v)\n"
        "assert g.info('device_cache_bytes') == 0\n"
        "g.destroy()\n"
        f"np.save({str(out)!r}, v)\n")
    env = dict(os.environ, LAGRBF_DEVICE_CACHE="0")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    print("case: device cache off in a child process", flush=True)
    r = subprocess.run([sys.executable, "-c", code], env=env, cwd=root, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    assert np.array_equal(np.load(out), here)
