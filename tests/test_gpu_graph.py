"""Captured realisations: a HIP-graph replay of the IC sequence for a staged
seed must equal an eager run with that seed, bit for bit, for every noise
plugin, across replays with eager work and synchronisation in between and
across streams."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    # n, box, LPT order, bulk precision, generator
    (32, 100.0, 3, "double", "MUSIC1"),
    (48, 150.0, 2, "float", "NGENIC"),
    (32, 120.0, 2, "float", "PANPHASIA_THREEFRY"),
]


def particles(ic):
    return [ic.particles(w) for w in range(6)]


@pytest.mark.parametrize("n,box,order,prec,gen", CASES)
def test_graph_replay_equals_eager(ctx, n, box, order, prec, gen):
    import monofonic_amd as M
    print(f"graph replay {gen} n={n} order={order} {prec}", flush=True)
    kw = dict(LPTorder=order, precision=prec, generator=gen, seed=11)
    eager = M.ICGenerator(ctx, n, box, 24.0, **kw)
    graph = M.ICGenerator(ctx, n, box, 24.0, **kw)
    graph.capture()
    ref = {}
    for s in (101, 202, 303):
        eager.set_seed(s)
        eager.run()
        ctx.synchronize()
        ref[s] = particles(eager)
    for r, s in enumerate((202, 101, 303, 202)):
        graph.set_seed(s)
        graph.launch()
        ctx.synchronize()
        got = particles(graph)
        for w in range(6):
            assert np.array_equal(got[w], ref[s][w]), (r, s, w)
        # eager work on the same stream between replays
        eager.set_seed(404 + r)
        eager.run()
        ctx.synchronize()


def test_graph_replays_on_two_streams(ctx):
    """two generators on two contexts (streams) replaying concurrently"""
    import monofonic_amd as M
    kw = dict(LPTorder=2, precision="float", generator="NGENIC")
    ctx2 = M.Context(0)
    ref = {}
    eager = M.ICGenerator(ctx, 48, 150.0, 24.0, seed=1, **kw)
    for s in (7, 8, 9, 10):
        eager.set_seed(s)
        eager.run()
        ctx.synchronize()
        ref[s] = particles(eager)
    # generators are addressed by index only: every reference to them must be
    # gone before their context is closed
    gens = [M.ICGenerator(c, 48, 150.0, 24.0, seed=1, **kw) for c in (ctx, ctx2)]
    for i in range(2):
        gens[i].capture()
    for rnd in range(2):
        seeds = (7 + 2 * rnd, 8 + 2 * rnd)
        for i in range(2):
            gens[i].set_seed(seeds[i])
        for i in range(2):
            gens[i].launch()
        ctx.synchronize()
        ctx2.synchronize()
        for i in range(2):
            got = particles(gens[i])
            for w in range(6):
                assert np.array_equal(got[w], ref[seeds[i]][w]), (rnd, seeds[i], w)
    gens.clear()
    ctx2.close()


def test_context_outlives_its_objects(ctx):
    """destroying a context that still has a generator on it is refused"""
    import monofonic_amd as M
    c2 = M.Context(0)
    gens = [M.ICGenerator(c2, 32, 100.0, 24.0, LPTorder=2, precision="float", generator="MUSIC1", seed=3)]
    with pytest.raises(M.MfxError, match="destroy them first"):
        c2.close()
    gens.clear()
    c2.close()


def test_batch_one_graph_equals_eager(ctx):
    """realisations of three grid sizes captured into one graph (branches on
    three streams): every generator's result equals an eager run with its
    seed, over two launches with new seeds"""
    import monofonic_amd as M
    kw = dict(LPTorder=2, precision="float", generator="NGENIC")
    sizes = [(32, 50.0), (48, 75.0), (64, 100.0)] * 2
    gens = [M.ICGenerator(ctx, n, L, 24.0, seed=1, **kw) for n, L in sizes]
    ref = [M.ICGenerator(ctx, n, L, 24.0, seed=1, **kw) for n, L in sizes[:3]]
    batch = M.Batch(gens, nstreams=3)
    for rnd in range(2):
        seeds = [500 + 10 * rnd + i for i in range(len(gens))]
        for g, s in zip(gens, seeds):
            g.set_seed(s)
        batch.launch()
        batch.synchronize()
        for i, (g, s) in enumerate(zip(gens, seeds)):
            r = ref[i % 3]
            r.set_seed(s)
            r.run()
            ctx.synchronize()
            for w in range(6):
                assert np.array_equal(g.particles(w), r.particles(w)), (rnd, i, w)
    # a generator in a batch cannot be destroyed before the batch
    with pytest.raises(M.MfxError, match="batch"):
        M._lib.call("mfx_icgen_destroy", gens[0]._h)
    batch.close()
    gens.clear()
    ref.clear()
