"""The host output boundary (include/mfx_io.h): files written by the generic
field writer (src/plugins/output_generic.cc:17-128, Grid_FFT::Write_to_HDF5
src/grid_fft.cc:669-760) and the Gadget-HDF5 particle writer
(src/plugins/output_gadget_hdf5.cc) read back with the HDF5 C library: names,
types, shapes, attributes and every value bit for bit."""
import numpy as np
import pytest

from hdf5_read import File


def test_generic_fields_from_an_oracle_realisation(orc, tmp_path):
    """a 32^3 2LPT realisation of the oracle pipeline written as the generic
    plugin writes it (DM_dx..DM_vz, displacements without the lattice, lunit =
    vunit = 1) and read back bit for bit"""
    from monofonic_amd.io import GenericOutput
    from test_gpu_lpt import lattice_q, oracle_pipeline
    n, box, z = 32, 100.0, 50.0
    cos = orc.Cosmology()
    fields, _, _ = oracle_pipeline(orc, n, box, z, 2, 4711, cos)
    names = ["DM_dx", "DM_dy", "DM_dz", "DM_vx", "DM_vy", "DM_vz"]
    data = [fields[w] - lattice_q(n, w) if w < 3 else fields[w] for w in range(6)]
    fn = tmp_path / "generic.hdf5"
    astart = 1 / (1 + z)
    with GenericOutput(fn, box, astart, 0.045, 0.3, 0.7) as out:
        for name, f in zip(names, data):
            padded = np.zeros((n, n, n + 2))
            padded[:, :, :n] = f
            out.write_field_padded(name, padded, n)
    with File(fn) as h:
        for key, val in (("Boxsize", box), ("astart", astart), ("Omega_b", 0.045), ("Omega_m", 0.3),
                         ("Omega_L", 0.7)):
            v, (dt, shape) = h.attr("Header", key)
            assert dt == np.float64 and shape is None and v == val, key
        for name, f in zip(names, data):
            got, dt = h.dataset(name)
            assert dt == np.float64 and got.shape == (n, n, n)
            assert np.array_equal(got, f), name


@pytest.mark.parametrize("prec", [np.float32, np.float64])
def test_generic_field_precision_and_extent(tmp_path, prec):
    from monofonic_amd.io import GenericOutput
    rng = np.random.default_rng(3)
    a = rng.standard_normal((6, 5, 8)).astype(prec)
    fn = tmp_path / "f.hdf5"
    with GenericOutput(fn, 1.0, 0.1, 0.0, 1.0, 0.0) as out:
        out.write_field("DM_delta", a)
    with File(fn) as h:
        got, dt = h.dataset("DM_delta")
    assert dt == np.dtype(prec) and np.array_equal(got, a)


@pytest.mark.parametrize("double_prec,long_ids,gadget2", [(True, True, False), (False, False, True)])
def test_gadget_particles(tmp_path, double_prec, long_ids, gadget2):
    from monofonic_amd.io import GadgetOutput
    n, box, z = 8, 50.0, 24.0
    rng = np.random.default_rng(9)
    pos = [rng.random(n ** 3) for _ in range(3)]
    vel = [rng.standard_normal(n ** 3) for _ in range(3)]
    a = 1 / (1 + z)
    lscale, vscale = box, box / np.sqrt(a)
    fn = tmp_path / "ics.hdf5"
    with GadgetOutput(fn, z, box, 0.3, 0.7, 0.045, 0.7, 3, n, double_prec=double_prec, long_ids=long_ids,
                      gadget2_compat=gadget2) as out:
        out.write_particles(pos, vel, n, 0.3, lscale, vscale)
    rt = np.float64 if double_prec else np.float32
    with File(fn) as h:
        c, dt = h.dataset("PartType1/Coordinates")
        assert dt == rt and c.shape == (n ** 3, 3)
        assert np.array_equal(c, np.stack([p * lscale for p in pos], 1).astype(rt))
        v, _ = h.dataset("PartType1/Velocities")
        assert np.array_equal(v, np.stack([x * vscale for x in vel], 1).astype(rt))
        ids, dt = h.dataset("PartType1/ParticleIDs")
        assert dt == (np.uint64 if long_ids else np.uint32)
        assert np.array_equal(ids, np.arange(n ** 3))
        assert not h.exists("PartType1/Masses")
        npt, (dt, shape) = h.attr("Header", "NumPart_ThisFile")
        assert shape == (6,) and dt == (np.uint32 if gadget2 else np.uint64)
        assert list(npt) == [0, n ** 3, 0, 0, 0, 0]
        mt, _ = h.attr("Header", "MassTable")
        assert mt[1] == 0.3 * (27.7519737 * box ** 3) / n ** 3 and mt[0] == 0
        for key, val in (("Time", a), ("Redshift", 1 / a - 1), ("BoxSize", box), ("Omega0", 0.3),
                         ("OmegaLambda", 0.7), ("HubbleParam", 0.7)):
            got, (dt, shape) = h.attr("Header", key)
            assert shape == (1,) and got[0] == val, key
        nf, _ = h.attr("Header", "NumFilesPerSnapshot")
        assert nf[0] == 1
        assert h.attr("ICs_parameters", "Code")[0] == "MUSIC2 - monofonIC"
        assert h.attr("ICs_parameters", "LPT Order")[0] == 3
        assert h.attr("ICs_parameters", "Grid Resolution")[0] == n
        assert h.attr("ICs_parameters", "Random Generator")[0] == "MUSIC1"
        if gadget2:
            assert list(h.attr("Header", "NumPart_Total_HighWord")[0]) == [0] * 6


@pytest.mark.parametrize("load,ov", [("bcc", 2), ("fcc", 4), ("rsc", 8)])
def test_gadget_particles_of_a_bravais_load(tmp_path, load, ov):
    """ov n^3 particles, sublattice-major; ParticleIDs = ov * cell + sublattice
    (particle_generator.hh:229-250); the particle mass divides by the
    particle count; the load is recorded in ICs_parameters"""
    from monofonic_amd.io import GadgetOutput
    n, box, z = 4, 50.0, 24.0
    rng = np.random.default_rng(ov)
    pos = [rng.random(ov * n ** 3) for _ in range(3)]
    vel = [rng.standard_normal(ov * n ** 3) for _ in range(3)]
    fn = tmp_path / "ics.hdf5"
    with GadgetOutput(fn, z, box, 0.3, 0.7, 0.045, 0.7, 2, n, double_prec=True, long_ids=True,
                      particle_load=load) as out:
        out.write_particles(pos, vel, n, 0.3, 1.0, 1.0)
        with pytest.raises(ValueError):  # n^3 values per component are not enough
            out.write_particles([p[: n ** 3] for p in pos], vel, n, 0.3, 1.0, 1.0)
    with File(fn) as h:
        c, _ = h.dataset("PartType1/Coordinates")
        assert c.shape == (ov * n ** 3, 3) and np.array_equal(c, np.stack(pos, 1))
        ids, _ = h.dataset("PartType1/ParticleIDs")
        p = np.arange(ov * n ** 3)
        assert np.array_equal(ids, ov * (p % n ** 3) + p // n ** 3)
        assert sorted(ids) == list(range(ov * n ** 3))
        npt, _ = h.attr("Header", "NumPart_Total")
        assert list(npt) == [0, ov * n ** 3, 0, 0, 0, 0]
        mt, _ = h.attr("Header", "MassTable")
        assert mt[1] == 0.3 * (27.7519737 * box ** 3) / (ov * n ** 3)
        assert h.attr("ICs_parameters", "Particle Load")[0] == load


def test_gadget_particles_with_baryons(tmp_path):
    """DoBaryons: both species with a Masses dataset ((1/N) Omega rhoc L^3,
    the mass table 0), the baryons (PartType0) numbered after the dark
    matter (IDoffset 1)"""
    from monofonic_amd.io import GadgetOutput
    n, box, z, ov = 4, 50.0, 24.0, 2
    rng = np.random.default_rng(3)
    arr = lambda: [rng.random(ov * n ** 3) for _ in range(3)]
    fn = tmp_path / "ics.hdf5"
    Om, Ob = 0.3, 0.045
    with GadgetOutput(fn, z, box, Om, 0.7, Ob, 0.7, 3, n, double_prec=True, long_ids=True, particle_load="bcc",
                      do_baryons=True) as out:
        out.write_particles(arr(), arr(), n, Om - Ob, 1.0, 1.0, species=1)
        out.write_particles(arr(), arr(), n, Ob, 1.0, 1.0, species=0)
    munit = 27.7519737 * box ** 3
    N = ov * n ** 3
    with File(fn) as h:
        mt, _ = h.attr("Header", "MassTable")
        assert list(mt) == [0.0] * 6
        npt, _ = h.attr("Header", "NumPart_Total")
        assert list(npt) == [N, N, 0, 0, 0, 0]
        for sp, om, off in ((1, Om - Ob, 0), (0, Ob, N)):
            m, dt = h.dataset(f"PartType{sp}/Masses")
            assert dt == np.float64 and m.shape == (N,)
            assert np.all(m == (1.0 / N) * (om * munit))
            ids, _ = h.dataset(f"PartType{sp}/ParticleIDs")
            p = np.arange(N)
            assert np.array_equal(ids, off + ov * (p % n ** 3) + p // n ** 3)
        assert h.attr("ICs_parameters", "Baryons")[0] == 1


def test_writer_reports_errors(tmp_path):
    from monofonic_amd.io import GenericOutput, MfxIoError
    with pytest.raises(MfxIoError, match="cannot create"):
        GenericOutput(tmp_path / "no" / "such" / "dir.hdf5", 1, 1, 0, 1, 0)
    with GenericOutput(tmp_path / "g.hdf5", 1, 1, 0, 1, 0) as out:
        out.write_field("DM_dx", np.zeros((2, 2, 2)))
        with pytest.raises(MfxIoError, match="DM_dx"):
            out.write_field("DM_dx", np.zeros((2, 2, 2)))
