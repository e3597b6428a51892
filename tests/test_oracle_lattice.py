"""Pins of the oracle's Bravais particle loads (Grid_FFT::shift_field,
grid_fft.hh:832-842, and set_positions/set_velocities of bcc/fcc/rsc,
particle_generator.hh:47-54,499-543,597-638): a field of plane waves below
the Nyquist frequency, shifted by s cells, equals the same waves evaluated at
the shifted points, i + s (closed form); the sublattice sequence reaches the
absolute offsets of each load; a planted sign error of the phase is caught."""
import numpy as np
import pytest

WAVES = [((1, 2, 3), 0.3, 1.0), ((-3, 0, 5), 1.1, 0.5), ((7, -6, 2), -0.4, 0.25), ((0, 0, 7), 0.0, 0.75)]


def waves(n, s=(0.0, 0.0, 0.0)):
    """sum of cosine waves sampled at (i + s_x, j + s_y, k + s_z), padded layout"""
    i = np.arange(n)
    I, J, K = np.meshgrid(i + s[0], i + s[1], i + s[2], indexing="ij")
    f = np.zeros((n, n, n + 2))
    for (a, b, c), ph, amp in WAVES:
        f[:, :, :n] += amp * np.cos(2 * np.pi * (a * I + b * J + c * K) / n + ph)
    return f


@pytest.mark.parametrize("s", [(0.5, 0.5, 0.5), (0.0, 0.5, 0.5), (0.25, -0.5, 0.0), (0.5, 0.0, 0.0)])
def test_shift_field_samples_waves_at_shifted_points(orc, s):
    n, box = 16, 100.0
    f = waves(n)
    orc.shift_field(f, box, s)
    ref = waves(n, s)
    assert np.max(np.abs(f[:, :, :n] - ref[:, :, :n])) < 1e-13


def test_shift_field_mutation_is_detected(orc):
    n, box = 16, 100.0
    try:
        orc.lattice_set_mutation(1)
        f = waves(n)
        orc.shift_field(f, box, (0.5, 0.0, 0.25))
        assert np.max(np.abs(f[:, :, :n] - waves(n, (0.5, 0.0, 0.25))[:, :, :n])) > 0.1
    finally:
        orc.lattice_set_mutation(0)


@pytest.mark.parametrize("lattice,nsub", [("sc", 1), ("bcc", 2), ("fcc", 4), ("rsc", 8)])
def test_lattice_values_reach_each_sublattice(orc, lattice, nsub):
    """set_velocities (field values) and set_positions ((i + s)/n lunit +
    field) of every sublattice, against the closed form"""
    n, box, lunit = 16, 100.0, 3.0
    f = waves(n)
    sh = orc.LATTICE_SHIFTS[lattice]
    assert len(sh) == nsub and tuple(sh[0]) == (0, 0, 0)
    vel = orc.lattice_values(f, box, lattice, None)
    assert vel.shape == (nsub, n, n, n)
    for idim in range(3):
        pos = orc.lattice_values(f, box, lattice, idim, lunit)
        for l, s in enumerate(sh):
            ref = waves(n, s)[:, :, :n]
            q = (np.arange(n) + s[idim]) / n * lunit
            shp = [1, 1, 1]
            shp[idim] = n
            assert np.max(np.abs(vel[l] - ref)) < 1e-13, (lattice, l)
            assert np.max(np.abs(pos[l] - q.reshape(shp) - ref)) < 1e-13, (lattice, l, idim)
    # the offsets are distinct points of the unit cell
    assert len({tuple(s) for s in sh}) == nsub
