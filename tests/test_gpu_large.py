"""The 512^3 (768^3 padded) and 1024^3 (1536^3 padded) FP32 configurations
through a property that does not depend on the grid size: for a band-limited
potential whose products stay below a third of the Nyquist wave number, the
Orszag 3/2 products are exact, so phi2, phi3 and A3 are the same continuous
fields at every resolution.  The large-grid fields sampled at every 8th
(512^3) or 16th (1024^3) point must equal the oracle's 64^3 FP64 fields
(ic_generator.cc:464-547)."""
import numpy as np
import pytest

BOX = 300.0
# (mode (mx, my, mz), amplitude): sums of up to three modes stay within
# n/3 = 21 of the 64^3 grid in every component
MODES = [((1, 2, 0), 1.0), ((0, 1, 3), 0.7), ((2, 0, 1), -0.5), ((1, -1, 2), 0.4)]


def kspace_phi(n, dtype):
    """k-space grid (padded layout) of phi(x) = sum_m a_m cos(2 pi m.x / n),
    unitary normalisation: coefficient a sqrt(n^3)/2 at m and -m"""
    a = np.zeros((n, n, n + 2), dtype=dtype)
    k = a.view(np.complex128 if dtype == np.float64 else np.complex64).reshape(n, n, n // 2 + 1)
    s = np.sqrt(float(n) ** 3) / 2
    for (mx, my, mz), amp in MODES:
        if mz > 0:
            k[mx % n, my % n, mz] += amp * s
        elif mz < 0:
            k[(-mx) % n, (-my) % n, -mz] += amp * s
        else:  # both +m and -m lie in the stored kz = 0 plane
            k[mx % n, my % n, 0] += amp * s
            k[(-mx) % n, (-my) % n, 0] += amp * s
    return a


def test_kspace_phi_is_the_cosine_sum(orc):
    n = 16
    a = kspace_phi(n, np.float64)
    orc.fft_c2r(a)
    x = np.arange(n)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    f = sum(amp * np.cos(2 * np.pi * (m[0] * X + m[1] * Y + m[2] * Z) / n) for m, amp in MODES)
    assert np.max(np.abs(a[:, :, :n] - f)) < 1e-13


@pytest.fixture(scope="module")
def band_limited_ref(orc):
    ns = 64
    r2, r3, rA = orc.lpt_terms(kspace_phi(ns, np.float64), BOX, 3)
    out = []
    for r in [r2, r3] + list(rA):
        ref = r.copy()
        orc.fft_c2r(ref)
        out.append(ref[:, :, :ns])
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("n,path", [(512, "streamed"), (512, "stored"), (1024, "streamed"), (1024, "auto")])
def test_lpt_sources_band_limited(ctx, band_limited_ref, n, path):
    import monofonic_amd as M
    ns = 64
    step = n // ns
    print(f"band-limited 3LPT sources at {n}^3 FP32 ({3 * n // 2}^3 padded), path {path}", flush=True)
    phi = M.Grid_FFT(ctx, n, BOX, "float", initialspace=M.kspace_id)
    phi.upload(kspace_phi(n, np.float32))
    out = [M.Grid_FFT(ctx, n, BOX, "float", initialspace=M.kspace_id) for _ in range(5)]
    conv = M.OrszagConvolver(ctx, n, BOX, "float")
    conv.set_lpt_path(path)
    conv.lpt_sources(phi, out[0], out[1], out[2:], 3)
    ctx.synchronize()
    print(f"  sources done on the {conv.lpt_path} path", flush=True)
    del conv
    names = ["phi2", "phi3", "A3x", "A3y", "A3z"]
    for g, ref, name in zip(out, band_limited_ref, names):
        g.FourierTransformBackward()
        got = g.rdata()[::step, ::step, ::step].astype(np.float64)
        g.reset()
        scale = np.sqrt(np.mean(ref ** 2))
        err = np.sqrt(np.mean((got - ref) ** 2)) / scale
        mx = np.max(np.abs(got - ref)) / scale
        print(f"  {name}: rms {err:.2e} max {mx:.2e} (field rms {scale:.3e})", flush=True)
        assert scale > 0
        assert err < 2e-5, name
        assert mx < 2e-4, name
