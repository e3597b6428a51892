"""Semiclassical (PPT) Eulerian fields on the GPU (generic_out_eulerian,
src/plugins/output_generic.cc:43,63-68; src/ic_generator.cc:674-790) against
the oracle's restatement: hbar, the density |psi|^2 - 1 and the velocities of
one realisation, graph replays, the generic HDF5 file, and the rejections.

Tolerances (differences relative to the field's RMS): FP64 runs agree with
the oracle to rounding, about 1e-14 RMS and 1e-13 maximum; FP32 runs to the
FP32 phase error of (phi + phi2)/(hbar D), which reaches tens of radians:
about 1e-6 RMS and 3e-5 maximum, the density's largest where |psi|^2 peaks."""
import numpy as np
import pytest

from conftest import rel_max, rel_rms
from hdf5_read import File

pytestmark = pytest.mark.gpu

BOX = 100.0


def oracle_ppt(orc, n, box, zstart, order, seed, generator="MUSIC1"):
    """noise, phi(1), phi(2) (LPTorder >= 2), growth scaling (g1 = -D,
    g2 = -3/7 D^2), c2r, then the semiclassical fields with
    vunit = D vfact / L (velocity_unit() = 1)"""
    from test_gpu_lpt import oracle_noise
    cos = orc.Cosmology()
    a = 1.0 / (1.0 + zstart)
    D, vfac = cos.Dplus(a), cos.vfact(a)
    wn = oracle_noise(orc, n, seed, generator)
    orc.noise_normalise(wn, (box / n / 2.0 / np.pi) ** 1.5)
    phi = np.zeros_like(wn)
    orc.phi1(phi, wn, box, cos.amplitude_table(n, box))
    phi2 = None
    if order > 1:
        phi2, _, _ = orc.lpt_terms(phi, box, 2)
        orc.scale(phi2, -3.0 / 7.0 * D * D)
        orc.fft_c2r(phi2)
    orc.scale(phi, -D)
    orc.fft_c2r(phi)
    return orc.ppt_fields(phi, phi2, box, D, D * vfac / box)


TOL = {"double": (1e-12, 1e-11, 1e-12), "float": (5e-5, 5e-4, 1e-6)}


def _check(ic, ref, prec, label):
    delta, v, hbar = ref
    tr, tm, th = TOL[prec]
    h = ic.hbar()
    print(f"  {label}: hbar {h:.6e} (oracle {hbar:.6e})", flush=True)
    assert abs(h / hbar - 1) < th
    got = [ic.density()] + [ic.particles(3 + d) for d in range(3)]
    want = [delta] + [v[d] for d in range(3)]
    for w, (g, r) in enumerate(zip(got, want)):
        print(f"  {label} field {w}: rms {rel_rms(g, r):.2e} max {rel_max(g, r):.2e}", flush=True)
        assert rel_rms(g, r) < tr, (label, w)
        assert rel_max(g, r) < tm, (label, w)
    # the drift formed structure: the density contrast is not small
    assert np.std(delta) > 0.02


@pytest.mark.parametrize("n,order,prec,gen,z", [(32, 2, "double", "MUSIC1", 24.0),
                                               (32, 1, "float", "PANPHASIA_THREEFRY", 20.0),
                                               (64, 3, "double", "NGENIC", 30.0),
                                               (48, 2, "float", "NGENIC", 24.0)])
def test_eulerian_fields_vs_oracle(ctx, orc, n, order, prec, gen, z):
    import monofonic_amd as M
    print(f"ppt n={n} order={order} {prec} {gen}", flush=True)
    ic = M.ICGenerator(ctx, n, BOX, z, LPTorder=order, precision=prec, generator=gen, seed=777,
                       field_output=True, generic_out_eulerian=True)
    ic.run()
    _check(ic, oracle_ppt(orc, n, BOX, z, order, 777, gen), prec, f"n={n}")


def test_eulerian_graph_replay_and_generic_file(ctx, orc, tmp_path):
    import monofonic_amd as M
    n, z = 32, 24.0
    print("ppt graph replay n=32", flush=True)
    ic = M.ICGenerator(ctx, n, BOX, z, LPTorder=2, precision="double", generator="MUSIC1", seed=5,
                       field_output=True, generic_out_eulerian=True)
    ic.capture()
    for seed in (21, 22):
        ic.set_seed(seed)
        ic.launch()
        ctx.synchronize()
        _check(ic, oracle_ppt(orc, n, BOX, z, 2, seed), "double", f"replay seed={seed}")
    ic.download_async([3, 4, 5, 6])
    got = ic.download_wait()
    fn = tmp_path / "eulerian.hdf5"
    ic.write_generic(fn)
    with File(fn) as h:
        for w, name in [(3, "DM_vx"), (4, "DM_vy"), (5, "DM_vz"), (6, "DM_delta")]:
            a, dt = h.dataset(name)
            assert dt == np.float64 and np.array_equal(a, got[w]), name
        for name in ("DM_dx", "DM_dy", "DM_dz"):
            with pytest.raises(KeyError):
                h.dataset(name)


def test_eulerian_rejections(ctx):
    import monofonic_amd as M
    print("ppt rejections", flush=True)
    with pytest.raises(M.MfxError, match="field output"):
        M.ICGenerator(ctx, 32, BOX, 24.0, LPTorder=2, generic_out_eulerian=True)
    with pytest.raises(M.MfxError, match="PLT"):
        M.ICGenerator(ctx, 32, BOX, 24.0, LPTorder=2, field_output=True, PLT=True, generic_out_eulerian=True)
    with pytest.raises(M.MfxError, match="grid size"):
        M.ICGenerator(ctx, 40, BOX, 24.0, LPTorder=2, field_output=True, generic_out_eulerian=True)
    ic = M.ICGenerator(ctx, 16, BOX, 24.0, LPTorder=1, field_output=True, generic_out_eulerian=True)
    ic.run()
    with pytest.raises(M.MfxError, match="displacement"):
        ic.particles(0)
    with pytest.raises(M.MfxError, match="displacement"):
        ic.download_async([0])
    plain = M.ICGenerator(ctx, 16, BOX, 24.0, LPTorder=1, field_output=True)
    with pytest.raises(M.MfxError, match="Eulerian"):
        plain.hbar()


def test_eulerian_256_mass_conservation(ctx):
    """At 256^3 FP32 (beyond the oracle's sizes): the drift is unitary and the
    phases have unit modulus, so sum |psi|^2 = n^3 and the mean density
    contrast vanishes to rounding (Parseval); hbar > 0 and the fields are
    finite."""
    import monofonic_amd as M
    n = 256
    print("ppt mass conservation n=256", flush=True)
    ic = M.ICGenerator(ctx, n, 500.0, 24.0, LPTorder=2, precision="float", generator="MUSIC1", seed=31,
                       field_output=True, generic_out_eulerian=True)
    ic.run()
    delta = ic.density()
    assert ic.hbar() > 0
    assert np.all(np.isfinite(delta)) and np.std(delta) > 0.02
    print(f"  mean delta {np.mean(delta):.3e}, std {np.std(delta):.3e}", flush=True)
    assert abs(np.mean(delta)) < 2e-6
    for d in range(3):
        assert np.all(np.isfinite(ic.particles(3 + d)))
