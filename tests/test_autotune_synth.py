"""Host-side planning and workload generation (no GPU)."""
import numpy as np

from musoc_solver_amd import autotune as A
from musoc_solver_amd import synth


def valid_ie(ie, j2, hmin, hmax):
    h = np.diff(np.concatenate([[1], ie]))
    return ie[-1] == j2 + 1 and (h >= 1).all() and (h[:-1] <= hmax).all() and (h >= hmin).all()


def test_adaptive_ie_is_valid_and_monotone_in_the_bound():
    case = synth.make_case(64, 40, 3, metric="latlon", lat_max=80.0, h_land=50.0)
    prev = None
    for g in (1e2, 1e4, 1e6, 1e9):
        ie = A.adaptive_ie(case.coefs, max_growth=g, hmax=8)
        assert valid_ie(ie, 40, 2, 8)
        if prev is not None:
            assert len(ie) <= prev
        prev = len(ie)


def test_adaptive_ie_takes_taller_blocks_on_a_smooth_operator():
    j2 = 40
    smooth = synth.ideal_operator(32, j2)
    rough = synth.make_case(32, j2, 4, metric="latlon", lat_max=80.0, h_land=50.0).coefs
    assert len(A.adaptive_ie(smooth, max_growth=1e4)) < len(A.adaptive_ie(rough, max_growth=1e4))


def test_reference_ie_rule():
    # NB0 = ANINT(J2/N0); IE(N) = MIN(1 + NN*N, J1); IE(NB0) = J1
    assert list(synth.reference_ie(96, 4)) == [25, 49, 73, 97]
    assert synth.nb0_for_height(720, 3) == 240
    ie = synth.reference_ie(50, synth.nb0_for_height(50, 3))
    assert ie[-1] == 51 and np.all(np.diff(ie) > 0)


def test_synthetic_operator_has_the_reference_form():
    case = synth.make_case(48, 24, 5, metric="latlon", lat_max=80.0, h_land=50.0)
    al, ab, ac, ar, at = [np.asarray(c, np.float64) for c in case.coefs]
    assert (al > 0).all() and (ar > 0).all()
    assert (ab[:, 0] == 0).all() and (at[:, -1] == 0).all()      # closed south/north faces
    assert (at[:, :-1] > 0).all()                                 # the march never divides by 0
    gamma = -(ac + al + ar + ab + at)
    assert np.allclose(gamma, gamma.flat[0], rtol=1e-4) and gamma.flat[0] > 0
    assert np.allclose(ar[:-1], al[1:], rtol=1e-6)                # shared east/west faces
    land = 1.0 - case.mask.mean()
    assert 0.25 < land < 0.35
    s = synth.make_rhs(case.mask, 6, 2)
    assert abs(s[:, :, 0][case.mask].mean()) < 1e-6 and (s[:, :, 0][~case.mask] == 0).all()


def test_basin_specs_cover_the_regional_size_range():
    """256 basins: sizes log-uniform in 64x64 .. 512x384 (I2 a multiple of 16,
    as the batched products need I2 % 4 == 0), land 20-45%, distinct seeds,
    reproducible from the seed."""
    sp = synth.basin_specs(256, 7000)
    assert sp == synth.basin_specs(256, 7000) and sp != synth.basin_specs(256, 7001)
    i2 = np.array([s[0] for s in sp])
    j2 = np.array([s[1] for s in sp])
    land = np.array([s[2] for s in sp])
    assert (i2 % 16 == 0).all() and i2.min() >= 64 and i2.max() <= 512
    assert (j2 % 8 == 0).all() and j2.min() >= 64 and j2.max() <= 384
    assert ((land >= 0.20) & (land <= 0.45)).all()
    assert len({s[3] for s in sp}) == 256
    # log-uniform: the geometric mean sits near the range's, both ends are reached
    assert abs(np.log(i2).mean() - np.log(np.sqrt(64 * 512))) < 0.15
    assert i2.min() <= 80 and i2.max() >= 448
