"""bcc, fcc and rsc particle loads of the IC driver (particle_generator.hh
:47-54,499-543,597-638, ic_generator.cc:131-139) and the baryon species on
the secondary lattice (DoBaryons, ic_generator.cc:153-168,608-671): every
sublattice's positions and velocities against the oracle, which runs the
reference's sequence of shift_field calls (grid_fft.hh:832-842) with its
transform round trips, while the library shifts the k-space components in
place."""
import numpy as np
import pytest

from conftest import rel_max, rel_rms

pytestmark = pytest.mark.gpu

BOX = 100.0


@pytest.mark.parametrize("lattice,n,order,prec,gen", [("bcc", 32, 3, "double", "MUSIC1"),
                                                      ("fcc", 32, 2, "double", "NGENIC"),
                                                      ("rsc", 32, 3, "double", "MUSIC1"),
                                                      ("bcc", 64, 3, "float", "NGENIC")])
def test_bravais_load_vs_oracle(ctx, orc, lattice, n, order, prec, gen):
    import monofonic_amd as M
    from test_gpu_lpt import oracle_pipeline
    print(f"lattice {lattice} n={n} order={order} {prec} {gen}", flush=True)
    ic = M.ICGenerator(ctx, n, BOX, 24.0, LPTorder=order, precision=prec, generator=gen, seed=321,
                       ParticleLoad=lattice)
    ic.run()
    ref, _, _ = oracle_pipeline(orc, n, BOX, 24.0, order, 321, orc.Cosmology(), generator=gen, lattice=lattice)
    nsub = len(orc.LATTICE_SHIFTS[lattice])
    tol = (1e-11, 1e-10) if prec == "double" else (1e-5, 1e-4)
    for w in range(6):
        got = ic.particles(w)
        assert got.shape == (nsub, n, n, n)
        for l, s in enumerate(orc.LATTICE_SHIFTS[lattice]):
            g, r = got[l], ref[w][l]
            if w < 3:  # displacements: subtract the sublattice's lattice coordinate
                q = ((np.arange(n) + s[w]) / n).reshape([n if d == w else 1 for d in range(3)])
                g, r = g - q, r - q
            assert rel_rms(g, r) < tol[0], (w, l)
            assert rel_max(g, r) < tol[1], (w, l)


def test_bravais_load_graph_replay_and_download(ctx):
    """a captured fcc realisation replays bit-equal to the eager run, and the
    asynchronous download copies every sublattice"""
    import monofonic_amd as M
    n = 32
    kw = dict(LPTorder=2, precision="float", generator="NGENIC", ParticleLoad="fcc")
    ic = M.ICGenerator(ctx, n, BOX, 24.0, seed=5, **kw)
    ic.set_seed(6)
    ic.run()
    eager = [ic.particles(w) for w in range(6)]
    g = M.ICGenerator(ctx, n, BOX, 24.0, seed=5, own_particles=True, **kw)
    g.capture()
    g.set_seed(6)
    g.launch()
    g.download_async(range(6))
    host = g.download_wait()
    for w in range(6):
        assert host[w].shape == (4, n, n, n)
        assert np.array_equal(host[w], eager[w]), w
    del ic, g


@pytest.mark.parametrize("lattice,prec", [("sc", "double"), ("bcc", "double"), ("fcc", "float")])
def test_baryons_on_the_secondary_lattice_vs_oracle(ctx, orc, lattice, prec):
    """DoBaryons: the dark matter on the primary lattice as without baryons,
    the baryons on the secondary lattice, both against the oracle"""
    import monofonic_amd as M
    from test_gpu_lpt import oracle_pipeline
    n, order = 32, 3
    print(f"baryons {lattice} {prec}", flush=True)
    ic = M.ICGenerator(ctx, n, BOX, 24.0, LPTorder=order, precision=prec, generator="MUSIC1", seed=77,
                       ParticleLoad=lattice, DoBaryons=True)
    ic.run()
    tol = (1e-11, 1e-10) if prec == "double" else (1e-5, 1e-4)
    for secondary, species in ((False, "dm"), (True, "baryon")):
        ref, _, _ = oracle_pipeline(orc, n, BOX, 24.0, order, 77, orc.Cosmology(), lattice=lattice,
                                    secondary=True) if secondary else \
            oracle_pipeline(orc, n, BOX, 24.0, order, 77, orc.Cosmology(), lattice=lattice)
        o = orc.SECOND_LATTICE_SHIFT[lattice] if secondary else (0, 0, 0)
        for w in range(6):
            got = ic.particles(w, species)
            r = ref[w] if got.ndim == 4 or lattice != "sc" or secondary else ref[w]
            got4 = got.reshape((-1, n, n, n))
            r4 = np.asarray(r).reshape((-1, n, n, n))
            for l, s in enumerate(orc.LATTICE_SHIFTS[lattice]):
                g, rr = got4[l], r4[l]
                if w < 3:
                    q = ((np.arange(n) + s[w] + o[w]) / n).reshape([n if d == w else 1 for d in range(3)])
                    g, rr = g - q, rr - q
                assert rel_rms(g, rr) < tol[0], (species, w, l)
                assert rel_max(g, rr) < tol[1], (species, w, l)


def test_baryon_gadget_file_from_device(ctx, tmp_path):
    """write_gadget of a DoBaryons bcc realisation: PartType1 (dark matter)
    and PartType0 (baryons) hold the downloaded arrays times the units"""
    import monofonic_amd as M
    from hdf5_read import File
    n = 16
    ic = M.ICGenerator(ctx, n, BOX, 24.0, LPTorder=2, precision="double", generator="NGENIC", seed=8,
                       ParticleLoad="bcc", DoBaryons=True, own_particles=True)
    ic.run()
    ic.download_async(list(range(6)) + list(range(7, 13)))
    host = ic.download_wait()
    fn = tmp_path / "ics.hdf5"
    ic.write_gadget(fn, double_prec=True, long_ids=True, arrays=host)
    a = 1 / (1 + 24.0)
    with File(fn) as h:
        for sp, base in ((1, 0), (0, 7)):
            c, _ = h.dataset(f"PartType{sp}/Coordinates")
            v, _ = h.dataset(f"PartType{sp}/Velocities")
            assert c.shape == (2 * n ** 3, 3)
            for d in range(3):
                assert np.array_equal(c[:, d], host[base + d].reshape(-1) * BOX)
                assert np.array_equal(v[:, d], host[base + 3 + d].reshape(-1) * (BOX / np.sqrt(a)))
            m, _ = h.dataset(f"PartType{sp}/Masses")
            assert m.shape == (2 * n ** 3,)
    del ic


def test_bravais_load_rejects_unsupported_combinations(ctx):
    import monofonic_amd as M
    for kw in (dict(PLT=True), dict(field_output=True), dict(assemble_each=True)):
        with pytest.raises(M.MfxError):
            M.ICGenerator(ctx, 32, BOX, 24.0, LPTorder=2, precision="double", ParticleLoad="bcc", **kw)
        with pytest.raises(M.MfxError):
            M.ICGenerator(ctx, 32, BOX, 24.0, LPTorder=2, precision="double", DoBaryons=True, **kw)
