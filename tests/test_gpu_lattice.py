"""bcc, fcc and rsc particle loads of the IC driver (particle_generator.hh
:47-54,499-543,597-638, ic_generator.cc:131-139): every sublattice's
positions and velocities against the oracle, which runs the reference's
sequence of shift_field calls (grid_fft.hh:832-842) with its transform round
trips, while the library shifts the k-space components in place."""
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


def test_bravais_load_rejects_unsupported_combinations(ctx):
    import monofonic_amd as M
    for kw in (dict(PLT=True), dict(field_output=True), dict(assemble_each=True)):
        with pytest.raises(M.MfxError):
            M.ICGenerator(ctx, 32, BOX, 24.0, LPTorder=2, precision="double", ParticleLoad="bcc", **kw)
