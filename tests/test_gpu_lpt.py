"""nLPT source terms and the full IC pipeline on the GPU against the oracle's
restatement of ic_generator::run (src/ic_generator.cc:249-944)."""
import numpy as np
import pytest

from conftest import rel_max, rel_rms
from helpers import random_kspace_field

pytestmark = pytest.mark.gpu

BOX = 100.0


def gpu_lpt_reference_sequence(M, conv, phi, phi2, phi3, A3, order):
    """The reference's convolve call sequence (ic_generator.cc:471-545), through
    the drop-in OrszagConvolver interface."""
    op = M.op
    conv.convolve_SumOfHessians(phi, (0, 0), phi, (1, 1), (2, 2), op.assign_to(phi2))
    conv.convolve_Hessians(phi, (1, 1), phi, (2, 2), op.add_to(phi2))
    conv.convolve_Hessians(phi, (0, 1), phi, (0, 1), op.subtract_from(phi2))
    conv.convolve_Hessians(phi, (0, 2), phi, (0, 2), op.subtract_from(phi2))
    conv.convolve_Hessians(phi, (1, 2), phi, (1, 2), op.subtract_from(phi2))
    phi2.apply_InverseLaplacian()
    if order < 3:
        return
    conv.convolve_Hessians(phi, (0, 0), phi, (1, 1), phi, (2, 2), op.assign_to(phi3))
    conv.convolve_Hessians(phi, (0, 1), phi, (0, 2), phi, (1, 2), op.multiply_add_to(phi3, 2.0))
    conv.convolve_Hessians(phi, (1, 2), phi, (1, 2), phi, (0, 0), op.subtract_from(phi3))
    conv.convolve_Hessians(phi, (0, 2), phi, (0, 2), phi, (1, 1), op.subtract_from(phi3))
    conv.convolve_Hessians(phi, (0, 1), phi, (0, 1), phi, (2, 2), op.subtract_from(phi3))
    conv.convolve_SumOfHessians(phi, (0, 0), phi2, (1, 1), (2, 2), op.multiply_add_to(phi3, -5.0 / 7.0))
    conv.convolve_SumOfHessians(phi, (1, 1), phi2, (2, 2), (0, 0), op.multiply_add_to(phi3, -5.0 / 7.0))
    conv.convolve_SumOfHessians(phi, (2, 2), phi2, (0, 0), (1, 1), op.multiply_add_to(phi3, -5.0 / 7.0))
    conv.convolve_Hessians(phi, (0, 1), phi2, (0, 1), op.multiply_add_to(phi3, +10.0 / 7.0))
    conv.convolve_Hessians(phi, (0, 2), phi2, (0, 2), op.multiply_add_to(phi3, +10.0 / 7.0))
    conv.convolve_Hessians(phi, (1, 2), phi2, (1, 2), op.multiply_add_to(phi3, +10.0 / 7.0))
    phi3.apply_InverseLaplacian()
    for idim in range(3):
        p, pp = (idim + 1) % 3, (idim + 2) % 3
        conv.convolve_Hessians(phi2, (idim, p), phi, (idim, pp), op.assign_to(A3[idim]))
        conv.convolve_Hessians(phi2, (idim, pp), phi, (idim, p), op.subtract_from(A3[idim]))
        conv.convolve_DifferenceOfHessians(phi, (p, pp), phi2, (p, p), (pp, pp), op.add_to(A3[idim]))
        conv.convolve_DifferenceOfHessians(phi2, (p, pp), phi, (p, p), (pp, pp), op.subtract_from(A3[idim]))
        A3[idim].apply_InverseLaplacian()


def _kgrid(M, ctx, n, prec):
    return M.Grid_FFT(ctx, n, BOX, prec, initialspace=M.kspace_id)


@pytest.mark.parametrize("n,order", [(32, 3), (64, 2)])
def test_lpt_sequence_vs_oracle(ctx, orc, n, order):
    import monofonic_amd as M
    print(f"lpt sequence n={n} order={order}", flush=True)
    phi_h = random_kspace_field(orc, n, 5, slope=-3.0)
    phi = _kgrid(M, ctx, n, "double")
    phi.upload(phi_h)
    phi2 = _kgrid(M, ctx, n, "double")
    phi3 = _kgrid(M, ctx, n, "double")
    A3 = [_kgrid(M, ctx, n, "double") for _ in range(3)]
    conv = M.OrszagConvolver(ctx, n, BOX, "double")
    gpu_lpt_reference_sequence(M, conv, phi, phi2, phi3, A3, order)
    r2, r3, rA = orc.lpt_terms(phi_h, BOX, order)
    assert rel_rms(phi2.kdata(), orc.kview(r2)) < 1e-12
    assert rel_max(phi2.kdata(), orc.kview(r2)) < 1e-10
    if order > 2:
        assert rel_rms(phi3.kdata(), orc.kview(r3)) < 1e-12
        assert rel_max(phi3.kdata(), orc.kview(r3)) < 1e-10
        for d in range(3):
            assert rel_rms(A3[d].kdata(), orc.kview(rA[d])) < 1e-12
            assert rel_max(A3[d].kdata(), orc.kview(rA[d])) < 1e-10


def oracle_noise(orc, n, seed, generator="MUSIC1", descriptor=None, cubesize=32):
    """the white-noise grid after Fill_Grid and wnoise.FourierTransformForward
    (ic_generator.cc:255-257): NGENIC and PANPHASIA fill k space directly, so
    the forward transform is a no-op"""
    if generator == "PANPHASIA":
        import re
        m = re.match(r"\[Panph6,L(\d+),\((\d+),(\d+),(\d+)\),S(\d+),(?:KK(\d+),)?CH", descriptor)
        lev, x, y, z, size, kk = m.groups()
        wn = np.zeros((n, n, n + 2))
        orc.kview(wn)[:] = orc.panphasia_noise((int(x), int(y), int(z)), int(lev), int(size), int(kk or 0), n)
        return wn
    if generator == "NGENIC":
        return orc.ngenic_fill(n, orc.ngenic_seedtable(n, seed))
    wn = orc.music1_fill(n, seed, cubesize) if generator == "MUSIC1" else orc.threefry_fill(n, seed)
    orc.fft_r2c(wn)
    return wn


def oracle_pipeline(orc, n, boxlen, zstart, order, seed, cos, plt_tabs=None, generator="MUSIC1", lattice="sc",
                    secondary=False, descriptor=None, cubesize=32, amp=None):
    """ic_generator::run restated on the CPU: white noise (MUSIC1, NGENIC or
    the Threefry stream), phi(1..3), A(3), growth scaling and FP64
    positions/velocities of the sc lattice, or, for lattice bcc/fcc/rsc,
    arrays (sublattice, n, n, n) as set_positions/set_velocities fill them
    (on the secondary, baryon lattice with secondary); with plt_tabs the
    assembly uses the PLT gradient (ENABLE_PLT builds)"""
    a = 1.0 / (1.0 + zstart)
    D = cos.Dplus(a)
    vfac = cos.vfact(a)
    volfac = (boxlen / n / 2.0 / np.pi) ** 1.5
    wn = oracle_noise(orc, n, seed, generator, descriptor, cubesize)
    orc.noise_normalise(wn, volfac)
    phi = np.zeros_like(wn)
    orc.phi1(phi, wn, boxlen, cos.amplitude_table(n, boxlen) if amp is None else amp)
    phi2 = phi3 = None
    A3 = (None, None, None)
    if order > 1:
        phi2, phi3, A3 = orc.lpt_terms(phi, boxlen, order)
    orc.scale(phi, -D)
    if order > 1:
        orc.scale(phi2, -3.0 / 7.0 * D * D)
    if order > 2:
        orc.scale(phi3, 1.0 / 3.0 * D ** 3)
        for x in A3:
            orc.scale(x, 1.0 / 7.0 * D ** 3)
    out = []
    for kind in range(2):
        for idim in range(3):
            if plt_tabs is None:
                f = orc.assemble_component(phi, phi2, phi3, A3, boxlen, order, kind, idim, vfac, 1.0)
            else:
                f = orc.assemble_component_plt(phi, phi2, phi3, A3, plt_tabs, boxlen, order, kind, idim, vfac, 1.0)
            if lattice != "sc" or secondary:
                out.append(orc.lattice_values(f, boxlen, lattice, idim if kind == 0 else None, 1.0, secondary))
            elif kind == 0:
                out.append(orc.set_positions(f, idim, 1.0).reshape(n, n, n))
            else:
                out.append(f[:, :, :n].copy())
    return out, D, vfac


def lattice_q(n, idim):
    q = np.arange(n) / n
    sh = [1, 1, 1]
    sh[idim] = n
    return np.broadcast_to(q.reshape(sh), (n, n, n))


@pytest.mark.parametrize("n,order,zstart", [(32, 3, 24.0), (64, 2, 50.0)])
def test_icgen_fp64_matches_oracle(ctx, orc, n, order, zstart):
    import monofonic_amd as M
    print(f"icgen fp64 n={n} order={order}", flush=True)
    ic = M.ICGenerator(ctx, n, BOX, zstart, LPTorder=order, precision="double", generator="MUSIC1", seed=4711,
                       fused=False)
    ic.run()
    cos = orc.Cosmology()
    ref, D, vfac = oracle_pipeline(orc, n, BOX, zstart, order, 4711, cos)
    dg, vg = ic.scalars()
    assert abs(dg / D - 1) < 1e-12 and abs(vg / vfac - 1) < 1e-12
    for w in range(6):
        got = ic.particles(w)
        if w < 3:
            psi_g = got - lattice_q(n, w)
            psi_r = ref[w] - lattice_q(n, w)
        else:
            psi_g, psi_r = got, ref[w]
        # tolerance relative to the field RMS (sigma_Psi)
        assert rel_rms(psi_g, psi_r) < 1e-11, w
        assert rel_max(psi_g, psi_r) < 1e-10, w


def test_icgen_music1_wrapping_cubes_vs_oracle(ctx, orc):
    """[random] cubesize 48 on a 64^3 grid: the cubes repeat and the zero-mean
    subtraction is not a constant shift (random_music_wnoise_generator.cc:44-65),
    so the IC sequence must carry it; FP64 2LPT against the oracle, and a graph
    replay with a second seed"""
    import monofonic_amd as M
    n, order, zstart = 64, 2, 50.0
    print(f"icgen fp64 n={n} MUSIC1 cubesize 48", flush=True)
    ic = M.ICGenerator(ctx, n, BOX, zstart, LPTorder=order, precision="double", generator="MUSIC1", seed=4711,
                       cubesize=48)
    cos = orc.Cosmology()
    ic.capture()
    for seed in (4711, 4712):
        ic.set_seed(seed)
        ic.launch()
        ctx.synchronize()
        ref, _, _ = oracle_pipeline(orc, n, BOX, zstart, order, seed, cos, cubesize=48)
        for w in range(6):
            got = ic.particles(w)
            if w < 3:
                psi_g, psi_r = got - lattice_q(n, w), ref[w] - lattice_q(n, w)
            else:
                psi_g, psi_r = got, ref[w]
            assert rel_rms(psi_g, psi_r) < 1e-11, (seed, w)
            assert rel_max(psi_g, psi_r) < 1e-10, (seed, w)
    # the wrapped noise differs from the default cube size's
    plain, _, _ = oracle_pipeline(orc, n, BOX, zstart, order, 4712, cos)
    assert rel_rms(ic.particles(3), plain[3]) > 0.1


def test_icgen_caller_amplitude_table(ctx, orc):
    """an amplitude table of another transfer function (here the
    Eisenstein-Hu table reshaped by a smooth factor): FP64 2LPT against the
    oracle run on the same table; 1LPT is linear in the table, so doubling it
    doubles the velocities bit for bit (scaling by 2 is exact)"""
    import monofonic_amd as M
    n, z = 32, 24.0
    print("icgen caller amplitude table n=32", flush=True)
    cos = orc.Cosmology()
    q = np.arange(3 * (n // 2) ** 2 + 1)
    amp = cos.amplitude_table(n, BOX) * (1.0 + 0.5 * np.sin(q / 40.0))
    ic = M.ICGenerator(ctx, n, BOX, z, LPTorder=2, precision="double", generator="MUSIC1", seed=77,
                       amplitude_table=amp)
    ic.run()
    ref, _, _ = oracle_pipeline(orc, n, BOX, z, 2, 77, cos, amp=amp)
    for w in range(6):
        got = ic.particles(w)
        psi_g, psi_r = (got - lattice_q(n, w), ref[w] - lattice_q(n, w)) if w < 3 else (got, ref[w])
        assert rel_rms(psi_g, psi_r) < 1e-11, w
        assert rel_max(psi_g, psi_r) < 1e-10, w
    a1 = M.ICGenerator(ctx, n, BOX, z, LPTorder=1, precision="double", generator="MUSIC1", seed=77,
                       amplitude_table=amp)
    a2 = M.ICGenerator(ctx, n, BOX, z, LPTorder=1, precision="double", generator="MUSIC1", seed=77,
                       amplitude_table=2.0 * amp)
    a1.run()
    a2.run()
    for w in range(3, 6):
        assert np.array_equal(a2.particles(w), 2.0 * a1.particles(w)), w
    with pytest.raises(M.MfxError, match="amplitude_table"):
        M.ICGenerator(ctx, n, BOX, z, LPTorder=1, amplitude_table=amp[:-1])


def test_icgen_fp32_bulk_vs_fp64(ctx):
    """FP32 FFT/convolution bulk with FP64 tables and assembly stays within
    1e-5 RMS of the FP64 path (displacements relative to sigma_Psi)."""
    import monofonic_amd as M
    n = 64
    res = {}
    for prec in ("double", "float"):
        ic = M.ICGenerator(ctx, n, BOX, 24.0, LPTorder=3, precision=prec, generator="MUSIC1", seed=99, fused=True)
        ic.run()
        res[prec] = [ic.particles(w) for w in range(6)]
    for w in range(6):
        a, b = res["float"][w], res["double"][w]
        if w < 3:
            a = a - lattice_q(n, w)
            b = b - lattice_q(n, w)
        assert rel_rms(a, b) < 1e-5, w


@pytest.mark.parametrize("n,order,path", [(32, 3, "streamed"), (32, 3, "stored"), (48, 3, "streamed"),
                                          (48, 3, "stored"), (64, 2, "auto")])
def test_fused_lpt_engine_vs_oracle(ctx, orc, n, order, path):
    """mfx_lpt_sources (Hessian-tensor engine) equals the reference's 5/33
    convolve calls, restated by the oracle, on both 3LPT paths"""
    import monofonic_amd as M
    print(f"fused lpt n={n} order={order} path={path}", flush=True)
    phi_h = random_kspace_field(orc, n, 8, slope=-3.0)
    phi = _kgrid(M, ctx, n, "double")
    phi.upload(phi_h)
    phi2, phi3 = _kgrid(M, ctx, n, "double"), _kgrid(M, ctx, n, "double")
    A3 = [_kgrid(M, ctx, n, "double") for _ in range(3)]
    conv = M.OrszagConvolver(ctx, n, BOX, "double")
    conv.set_lpt_path(path)
    conv.lpt_sources(phi, phi2, phi3, A3, order)
    assert conv.lpt_path == (path if order > 2 else "auto")
    r2, r3, rA = orc.lpt_terms(phi_h, BOX, order)
    pairs = [(phi2, r2)] + ([(phi3, r3)] + list(zip(A3, rA)) if order > 2 else [])
    for g, r in pairs:
        assert rel_rms(g.kdata(), orc.kview(r)) < 1e-12
        assert rel_max(g.kdata(), orc.kview(r)) < 1e-10


@pytest.mark.parametrize("n", [64, 128])
def test_fused_fp32_vs_reference_sequence_fp32(ctx, orc, n):
    """FP32: both 3LPT paths of the fused engine against the reference's
    convolve sequence run in FP32 on the same device"""
    import monofonic_amd as M
    print(f"fused fp32 paths n={n}", flush=True)
    phi_h = random_kspace_field(orc, n, 9, slope=-3.0)
    res = {}
    for mode in ("streamed", "stored", "sequence"):
        phi = _kgrid(M, ctx, n, "float")
        phi.upload(phi_h.astype(np.float32))
        out = [_kgrid(M, ctx, n, "float") for _ in range(5)]
        conv = M.OrszagConvolver(ctx, n, BOX, "float")
        if mode == "sequence":
            gpu_lpt_reference_sequence(M, conv, phi, out[0], out[1], out[2:], 3)
        else:
            conv.set_lpt_path(mode)
            conv.lpt_sources(phi, out[0], out[1], out[2:], 3)
        res[mode] = [g.kdata() for g in out]
    for mode in ("streamed", "stored"):
        for a, b in zip(res[mode], res["sequence"]):
            assert rel_rms(a, b) < 1e-5, mode


@pytest.mark.parametrize("fused,each", [(True, False), (True, True), (False, False)])
def test_icgen_fused_matches_oracle(ctx, orc, fused, each):
    """both LPT engines, and the six-component and one-at-a-time assembly"""
    import monofonic_amd as M
    n = 32
    ic = M.ICGenerator(ctx, n, BOX, 24.0, LPTorder=3, precision="double", generator="MUSIC1", seed=31, fused=fused,
                       assemble_each=each)
    ic.run()
    ref, _, _ = oracle_pipeline(orc, n, BOX, 24.0, 3, 31, orc.Cosmology())
    for w in range(6):
        a = ic.particles(w) - (lattice_q(n, w) if w < 3 else 0)
        b = ref[w] - (lattice_q(n, w) if w < 3 else 0)
        assert rel_rms(a, b) < 1e-11, w


@pytest.mark.parametrize("prec,path", [("double", "stored"), ("double", "streamed"), ("float", "stored")])
def test_vorticity_invariant(ctx, orc, prec, path):
    """The Lagrangian vorticity invariant of src/testing.cc:102-241 from the
    library's LPT sources, with its products formed by the library's own
    convolve_Gradients as the reference does: it falls as D^4 (x16 when D is
    halved) with A3 and as D^3 (x8) without (lpt_invariant.py)"""
    import monofonic_amd as M
    from lpt_invariant import band_limited_phi, displacement_velocity, grads
    n = 64
    print(f"vorticity invariant {prec} {path}", flush=True)
    cdt = np.complex128 if prec == "double" else np.complex64
    a = band_limited_phi(orc, n, 7, BOX, n // 8)
    phi = _kgrid(M, ctx, n, prec)
    phi.upload(a)
    phi2, phi3 = _kgrid(M, ctx, n, prec), _kgrid(M, ctx, n, prec)
    A3 = [_kgrid(M, ctx, n, prec) for _ in range(3)]
    conv = M.OrszagConvolver(ctx, n, BOX, prec)
    conv.set_lpt_path(path)
    conv.lpt_sources(phi, phi2, phi3, A3, 3)
    fields = [g.kdata().astype(np.complex128) for g in (phi, phi2, phi3)]
    fields.append([g.kdata().astype(np.complex128) for g in A3])
    G = grads(n, BOX)

    def kgrid_of(kc):
        g = _kgrid(M, ctx, n, prec)
        g.upload(np.ascontiguousarray(kc.astype(cdt)).view(g.dtype).reshape(n, n, n + 2))
        return g

    def rms(D, with_a3):
        psi, vel = displacement_velocity(G, *fields, D, with_a3)
        gp, gv = [kgrid_of(x) for x in psi], [kgrid_of(x) for x in vel]
        out = []
        for i in range(3):
            j = (i + 1) % 3
            inv = kgrid_of(G[j] * vel[i] - G[i] * vel[j])
            for k in range(3):
                conv.convolve_Gradients(gp[k], (i,), gv[k], (j,), M.op.add_to(inv))
                conv.convolve_Gradients(gp[k], (j,), gv[k], (i,), M.op.subtract_from(inv))
            inv.FourierTransformBackward()
            out.append(np.sqrt(np.mean(inv.rdata().astype(np.float64) ** 2)))
        return np.array(out)

    w = [rms(D, True) for D in (1.0, 0.5)]
    wo = [rms(D, False) for D in (1.0, 0.5)]
    print(f"  with A3 {w[0]} / {w[1]}, without {wo[0]} / {wo[1]}", flush=True)
    assert np.all(np.abs(w[0] / w[1] - 16) < 0.5), w
    assert np.all(np.abs(wo[0] / wo[1] - 8) < 0.3), wo
    assert np.all(w[1] < wo[1] / 20)
