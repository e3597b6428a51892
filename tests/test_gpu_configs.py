"""The workload configurations at their real transform sizes, through
ICGenerator (src/ic_generator.cc:249-944) against the FP64 oracle:

  3LPT FP64 256^3 (384^3 padded), box 500, z 24, MUSIC1: the whole pipeline
      against the oracle (FP64 1e-11 rms / 1e-10 max relative to the field
      RMS; FP32 bulk 1e-5 rms)
  batched 2LPT 96^3 / 128^3 / 192^3, NGENIC: 96^3 against the oracle; 128^3
      and 192^3 FP32 HIP-graph replays bit-equal to eager runs
  1024^3 + PLT + Threefry: the Threefry fill sampled at 1024^3 against the
      oracle's cells, the PLT lattice gradient at 1024^3 modes
  a graph replay compared with the oracle directly

The 512^3 (768^3 padded) FP32 LPT sources are covered by test_gpu_large.py.
"""
import numpy as np
import pytest

from conftest import rel_max, rel_rms

pytestmark = pytest.mark.gpu

SEED = 271828


def lattice_q(n, idim):
    q = np.arange(n) / n
    sh = [1, 1, 1]
    sh[idim] = n
    return np.broadcast_to(q.reshape(sh), (n, n, n))


def compare_fields(ic, ref, n, rms_tol, max_tol, label):
    worst = 0.0
    for w in range(6):
        got = ic.particles(w)
        if w < 3:
            a = got - lattice_q(n, w)
            b = ref[w] - lattice_q(n, w)
        else:
            a, b = got, ref[w]
        r, m = rel_rms(a, b), rel_max(a, b)
        print(f"  {label} field {w}: rms {r:.2e} max {m:.2e}", flush=True)
        worst = max(worst, r)
        assert r < rms_tol, (label, w, r)
        assert m < max_tol, (label, w, m)
        del got, a, b
    return worst


@pytest.fixture(scope="module")
def c2_ref(orc):
    from test_gpu_lpt import oracle_pipeline
    print("oracle: 3LPT 256^3 FP64, box 500, z 24, MUSIC1", flush=True)
    ref, D, vfac = oracle_pipeline(orc, 256, 500.0, 24.0, 3, SEED, orc.Cosmology())
    print("oracle done", flush=True)
    return ref


@pytest.mark.parametrize("prec,rms_tol,max_tol", [("double", 1e-11, 1e-10), ("float", 1e-5, 3e-4)])
def test_3lpt_256_vs_oracle(ctx, c2_ref, prec, rms_tol, max_tol):
    """3LPT at 256^3 (padded 384^3: column length 384, z rows of 192 complex,
    the fused zA/zC kernels at K = 192) against the oracle"""
    import monofonic_amd as M
    print(f"icgen 3LPT 256^3 {prec}", flush=True)
    ic = M.ICGenerator(ctx, 256, 500.0, 24.0, LPTorder=3, precision=prec, generator="MUSIC1", seed=SEED)
    ic.run()
    compare_fields(ic, c2_ref, 256, rms_tol, max_tol, f"256^3 {prec}")


@pytest.mark.parametrize("prec,rms_tol,max_tol", [("double", 1e-11, 1e-10), ("float", 1e-5, 3e-4)])
def test_2lpt_96_ngenic_vs_oracle(ctx, orc, prec, rms_tol, max_tol):
    """one realisation of the batched class 96^3 (box 150, padded 144^3)"""
    import monofonic_amd as M
    from test_gpu_lpt import oracle_pipeline
    n, box, seed = 96, 150.0, 1003
    print(f"icgen 2LPT 96^3 NGENIC {prec}", flush=True)
    ic = M.ICGenerator(ctx, n, box, 24.0, LPTorder=2, precision=prec, generator="NGENIC", seed=seed)
    ic.run()
    ref, _, _ = oracle_pipeline(orc, n, box, 24.0, 2, seed, orc.Cosmology(), generator="NGENIC")
    compare_fields(ic, ref, n, rms_tol, max_tol, f"96^3 {prec}")


@pytest.mark.parametrize("n,box", [(128, 200.0), (192, 300.0)])
def test_batched_class_graph_equals_eager(ctx, n, box):
    """128^3 and 192^3 2LPT FP32 NGENIC realisations: a captured replay for a
    staged seed equals an eager run with that seed bit for bit"""
    import monofonic_amd as M
    print(f"graph vs eager 2LPT {n}^3 NGENIC FP32", flush=True)
    kw = dict(LPTorder=2, precision="float", generator="NGENIC", seed=5)
    eager = M.ICGenerator(ctx, n, box, 24.0, **kw)
    graph = M.ICGenerator(ctx, n, box, 24.0, **kw)
    graph.capture()
    for s in (2001, 2002):
        eager.set_seed(s)
        eager.run()
        graph.set_seed(s)
        graph.launch()
        ctx.synchronize()
        for w in range(6):
            assert np.array_equal(graph.particles(w), eager.particles(w)), (n, s, w)


# MUSIC1 needs a power-of-two grid (random_music.cc:69-74)
@pytest.mark.parametrize("gen,n", [("MUSIC1", 64), ("NGENIC", 48)])
def test_graph_replay_vs_oracle(ctx, orc, gen, n):
    """a HIP-graph replay for a staged seed against the oracle pipeline with
    that seed"""
    import monofonic_amd as M
    from test_gpu_lpt import oracle_pipeline
    box = 150.0
    print(f"graph replay vs oracle 3LPT {n}^3 {gen} FP64", flush=True)
    ic = M.ICGenerator(ctx, n, box, 24.0, LPTorder=3, precision="double", generator=gen, seed=1)
    ic.capture()
    for s in (77, 78):
        ic.set_seed(s)
        ic.launch()
        ctx.synchronize()
        ref, _, _ = oracle_pipeline(orc, n, box, 24.0, 3, s, orc.Cosmology(), generator=gen)
        compare_fields(ic, ref, n, 1e-11, 1e-10, f"replay seed {s}")


def test_threefry_fill_1024_sampled(ctx, orc):
    """the Threefry white noise of the 1024^3 configuration (FP32 grid)
    against the oracle's cells: level l = 10, every corner and edge of the
    grid plus random cells"""
    import monofonic_amd as M
    n, seed = 1024, 4242
    print("Threefry fill 1024^3 FP32", flush=True)
    g = M.Grid_FFT(ctx, n, 2000.0, "float")
    M.RNG_threefry(seed).Fill_Grid(g)
    a = g.download()
    g.reset()
    rng = np.random.default_rng(1)
    ijk = rng.integers(0, n, (4000, 3))
    corners = np.array([[i, j, k] for i in (0, n - 1) for j in (0, n - 1) for k in (0, n - 1)])
    ijk = np.concatenate([ijk, corners, [[512, 511, 1], [1, 1023, 512]]])
    got = a[ijk[:, 0], ijk[:, 1], ijk[:, 2]].astype(np.float64)
    ref = orc.threefry_cells(n, seed, ijk).astype(np.float32).astype(np.float64)
    # the uniforms are bit-exact (test_gpu_rng); log/cos of the Box-Muller
    # transform may differ in the last double ulp, which the rounding to
    # float almost never shows: at most one float ulp
    ulp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
    assert np.all(np.abs(got - ref) <= ulp), np.max(np.abs(got - ref) / ulp)
    assert np.all(a[:, :, n:] == 0)
    # unit variance over the whole grid
    s = float(np.std(a[:: 4, :: 4, :n].astype(np.float64)))
    assert abs(s - 1) < 1e-3, s


def test_plt_gradient_1024(ctx, orc):
    """the PLT lattice gradient (particle_plt.hh:551-585) at modes of the
    1024^3 grid, box 2000: CIC interpolation of the 64^3 tables at
    ijk * 64/1024 (quarter-sixteenth offsets) and the spherical projection,
    against the oracle on the same tables"""
    import monofonic_amd as M
    n, L = 1024, 2000.0
    plt = M.PLT(ctx, 64)
    tabs = plt.tables()
    rng = np.random.default_rng(9)
    ijk = np.stack([rng.integers(0, n, 600), rng.integers(0, n, 600), rng.integers(0, n // 2 + 1, 600)], axis=1)
    ijk = np.concatenate([ijk, [[0, 0, 1], [n // 2, n // 2, n // 2], [n - 1, 1, n // 2], [1, 0, 0], [17, 1000, 3]]])
    g, v = plt.gradient(n, L, ijk)
    for t in range(len(ijk)):
        gr, vr = orc.plt_gradient(tabs, n, L, *[int(x) for x in ijk[t]])
        scale = max(np.linalg.norm(gr), 1e-300)
        assert np.max(np.abs(g[t] - gr)) <= 1e-13 * scale, ijk[t]
        assert abs(v[t] - vr) < 1e-14, ijk[t]
