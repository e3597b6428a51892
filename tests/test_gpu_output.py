"""The output boundary on the device side: field output (the generic
plugin's DM_dx..DM_vz and DM_delta, src/plugins/output_generic.cc,
src/ic_generator.cc:926-955) against the oracle, asynchronous downloads that
overlap the next realisation, and HDF5 files written from them read back bit
for bit."""
import numpy as np
import pytest

from conftest import rel_max, rel_rms
from hdf5_read import File

pytestmark = pytest.mark.gpu


def oracle_fields(orc, n, box, z, order, seed):
    """the oracle pipeline's displacements (without the lattice), velocities
    and the 1st-order density c2r(D |k|^2 phi)"""
    from test_gpu_lpt import lattice_q, oracle_noise, oracle_pipeline
    cos = orc.Cosmology()
    out, D, _ = oracle_pipeline(orc, n, box, z, order, seed, cos)
    fields = [out[w] - lattice_q(n, w) if w < 3 else out[w] for w in range(6)]
    wn = oracle_noise(orc, n, seed)
    orc.noise_normalise(wn, (box / n / 2.0 / np.pi) ** 1.5)
    phi = np.zeros_like(wn)
    orc.phi1(phi, wn, box, cos.amplitude_table(n, box))
    k = np.fft.fftfreq(n, 1.0 / n) * 2 * np.pi / box
    kz = np.arange(n // 2 + 1) * 2 * np.pi / box
    k2 = k[:, None, None] ** 2 + k[None, :, None] ** 2 + kz[None, None, :] ** 2
    pk = orc.kview(phi)
    pk *= D * k2
    orc.fft_c2r(phi)
    fields.append(phi[:, :, :n].copy())
    return fields


def test_field_output_vs_oracle_and_generic_file(ctx, orc, tmp_path):
    import monofonic_amd as M
    n, box, z, seed = 32, 100.0, 50.0, 4711
    ic = M.ICGenerator(ctx, n, box, z, LPTorder=2, precision="double", generator="MUSIC1", seed=seed,
                       field_output=True)
    ic.run()
    ic.download_async(range(7))
    got = ic.download_wait()
    ref = oracle_fields(orc, n, box, z, 2, seed)
    for w in range(7):
        print(f"  field {w}: rms {rel_rms(got[w], ref[w]):.2e}", flush=True)
        assert rel_rms(got[w], ref[w]) < 1e-11, w
        assert rel_max(got[w], ref[w]) < 1e-10, w
    fn = tmp_path / "generic.hdf5"
    ic.write_generic(fn)
    names = ["DM_dx", "DM_dy", "DM_dz", "DM_vx", "DM_vy", "DM_vz", "DM_delta"]
    with File(fn) as h:
        for w, name in enumerate(names):
            a, dt = h.dataset(name)
            assert dt == np.float64 and np.array_equal(a, got[w]), name
        assert h.attr("Header", "Boxsize")[0] == box


def test_download_overlaps_next_realisation(ctx):
    """own_particles: the download of realisation 1 runs while realisation 2
    is computed; the copies hold realisation 1, the device realisation 2"""
    import monofonic_amd as M
    n, box = 64, 150.0
    kw = dict(LPTorder=3, precision="float", generator="NGENIC")
    ref = M.ICGenerator(ctx, n, box, 24.0, seed=1, **kw)
    ic = M.ICGenerator(ctx, n, box, 24.0, seed=1, own_particles=True, **kw)
    expect = {}
    for s in (11, 12):
        ref.set_seed(s)
        ref.run()
        expect[s] = [ref.particles(w) for w in range(6)]
    ic.set_seed(11)
    ic.run()
    ic.download_async()
    ic.set_seed(12)
    ic.run()
    host = ic.download_wait()
    ctx.synchronize()
    for w in range(6):
        assert np.array_equal(host[w], expect[11][w]), w
        assert np.array_equal(ic.particles(w), expect[12][w]), w


@pytest.mark.parametrize("double_prec", [True, False])
def test_gadget_file_from_device_particles(ctx, tmp_path, double_prec):
    import monofonic_amd as M
    n, box, z = 32, 100.0, 24.0
    ic = M.ICGenerator(ctx, n, box, z, LPTorder=3, precision="double", generator="MUSIC1", seed=3)
    ic.run()
    ic.download_async()
    host = [x.copy() for x in ic.download_wait()[:6]]
    fn = tmp_path / "ics.hdf5"
    ic.write_gadget(fn, double_prec=double_prec, arrays=host)
    a = 1 / (1 + z)
    rt = np.float64 if double_prec else np.float32
    with File(fn) as h:
        c, dt = h.dataset("PartType1/Coordinates")
        assert dt == rt
        assert np.array_equal(c, np.stack([x.ravel() * box for x in host[:3]], 1).astype(rt))
        v, _ = h.dataset("PartType1/Velocities")
        assert np.array_equal(v, np.stack([x.ravel() * (box / np.sqrt(a)) for x in host[3:]], 1).astype(rt))
        assert h.attr("Header", "NumPart_Total")[0][1] == n ** 3
