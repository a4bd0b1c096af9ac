"""Pins of the oracle's building blocks by what the mathematics fixes (CPU only).

Each check is paired with a planted mistake in the oracle (the plant build,
`oracle.planted`) that it must catch, so a check that a plausible slip would
pass does not count as a pin:

* xoshiro256++ (rng.hpp:59-85): outputs from state (1, 2, 3, 4) derived by
  hand, and the GF(2) period 2^256 - 1 of the state update;
* SamplingTable::Sample / DataTable::Sample (table.cpp:50-66, 100-118): the
  piecewise-linear inverse CDF and interpolant (numpy.interp), end clamps,
  flat CDF segments;
* the azimuth rejection sampler (phase.cpp:27-49): the exact CDF of F and
  the mean trial count Fmax / mean(F);
* HG and isotropic theta (phase.cpp:78-83, 108-115): CDF(cos theta(u)) = u,
  mean cosine g;
* the frame and Jones step (simulation.cpp:438-492): the spherical angles of
  the new direction in the old frame, orthonormal right-handed frames, the
  amplitude matrix acting on the parallel/perpendicular field components;
* free path and interface rescale (medium.cpp:123-126, simulation.cpp:317-383):
  ballistic transmission exp(-z/mfp), and through two layers with the
  remaining optical depth carried over; a stack of identical layers equals one
  layer on the same streams;
* isotropic scalar transport: total diffuse reflectance of a half-space at
  normal incidence, 1 - H(1) sqrt(1 - a) (Chandrasekhar's H function);
* sources (laser.cpp:11-75): spatial and temporal moments;
* planar fluence (detector.cpp:780-841): spot centred on an off-axis source,
  Stokes of the lab field with x negated; planar field (:546-568): phase
  k0 * n * path;
* postprocess_farfield_cbs (results.cpp:61-103): solid angles and scaling.
"""
import math

import numpy as np
import pytest

import luminis_mc as lmc
import oracle
import rng_check
import scenes
from luminis_mc import _lib
from luminis_mc.simulation import SceneBuild

WL = scenes.WAVELENGTH


# --------------------------------------------------------------- xoshiro ----
def test_xoshiro256pp_first_outputs_by_hand():
    """State (1, 2, 3, 4): out = rotl(s0 + s3, 23) + s0 = 5 * 2^23 + 1 = 41943041;
    t = s1 << 17 = 262144; s2 ^= s0 -> 2; s3 ^= s1 -> 6; s1 ^= s2 -> 0;
    s0 ^= s3 -> 7; s2 ^= t -> 262146; s3 = rotl(6, 45) = 3 * 2^46.  Next
    output rotl(7 + 3 * 2^46, 23) + 7 = (7 * 2^23 + 3 * 2^5) + 7 = 58720359
    (bits 46, 47 rotate to 5, 6)."""
    out, st = oracle.xoshiro_state_seq((1, 2, 3, 4), 2)
    assert [int(v) for v in out] == [41943041, 58720359]
    assert [int(v) for v in st[0]] == [7, 0, 262146, 3 * 2 ** 46]
    with oracle.planted("xo_rot"):
        out, st = oracle.xoshiro_state_seq((1, 2, 3, 4), 2)
        assert [int(v) for v in st[0]] != [7, 0, 262146, 3 * 2 ** 46]


def test_xoshiro256pp_state_update_has_full_period():
    """The state update is linear over GF(2); it has period 2^256 - 1 when the
    minimal polynomial of one state bit is primitive of degree 256.  The
    factor list of 2^256 - 1 (Fermat numbers F0..F7) is checked here too."""
    primes = rng_check.PRIMES_2_256_MINUS_1
    assert math.prod(primes) == (1 << 256) - 1 and all(rng_check.is_probable_prime(q) for q in primes)
    _, st = oracle.xoshiro_state_seq((0x0123456789ABCDEF, 0xFEDCBA9876543210, 0x0F1E2D3C4B5A6978, 1), 600)
    bits = [int(v) & 1 for v in st[:, 0]]
    assert rng_check.sequence_is_primitive(bits, 256, primes)
    with oracle.planted("xo_rot"):
        _, st = oracle.xoshiro_state_seq((0x0123456789ABCDEF, 0xFEDCBA9876543210, 0x0F1E2D3C4B5A6978, 1), 600)
        assert not rng_check.sequence_is_primitive([int(v) & 1 for v in st[:, 0]], 256, primes)


# ---------------------------------------------------------------- tables ----
def _tables_case(rng, n):
    pdf = rng.uniform(0.1, 2.0, n)
    cdf = np.cumsum(pdf)
    return cdf / cdf[-1]


def test_table_sample_is_the_linear_inverse_cdf():
    rng = np.random.default_rng(1)
    for n in (2, 3, 17, 200):
        cdf = _tables_case(rng, n)
        x0, st = -0.3, 0.25
        nodes = x0 + st * np.arange(n)
        u = np.concatenate([rng.uniform(0, 1, 300), cdf, [0.0, cdf[0] / 2, 1.0, 1.5]])
        got = oracle.table_sample(cdf, x0, st, u)
        assert np.allclose(got, np.interp(u, cdf, nodes), rtol=0, atol=1e-14), n
    # a flat CDF segment (zero pdf): u on the plateau maps to its left end,
    # just above it to its right end (lower_bound picks the first node >= u)
    cdf = np.array([0.0, 0.5, 0.5, 1.0])
    got = oracle.table_sample(cdf, 0.0, 1.0, [0.0, 0.25, 0.5, 0.5 + 1e-12, 0.75, 1.0])
    assert np.allclose(got, [0.0, 0.5, 1.0, 2.0, 2.5, 3.0], rtol=0, atol=1e-9)
    with oracle.planted("table_t"):
        cdf = _tables_case(np.random.default_rng(1), 17)
        u = np.random.default_rng(2).uniform(0, 1, 200)
        assert not np.allclose(oracle.table_sample(cdf, 0.0, 1.0, u), np.interp(u, cdf, np.arange(17.0)), atol=1e-6)


def test_data_sample_is_linear_interpolation_with_end_clamps():
    rng = np.random.default_rng(3)
    for n in (2, 5, 1000):
        y = rng.normal(size=n) + 1j * rng.normal(size=n)
        dx = math.pi / max(n - 1, 1)
        grid = dx * np.arange(n)
        x = np.concatenate([rng.uniform(-0.2, grid[-1] + 0.2, 400), grid, [-1.0, 0.0, grid[-1], 10.0]])
        got = oracle.data_sample(y, dx, x)
        ref = np.interp(x, grid, y.real) + 1j * np.interp(x, grid, y.imag)
        # the position x * (1/dx) carries ~n ulps (the reference multiplies by 1/dx too)
        tol = 1e-15 + 4e-16 * n * np.abs(np.diff(y)).max(initial=0.0)
        assert np.abs(got - ref).max() < tol, n
    with oracle.planted("data_t"):
        y = rng.normal(size=9) + 0j
        x = rng.uniform(0, 8 * 0.1, 100)
        assert not np.allclose(oracle.data_sample(y, 0.1, x), np.interp(x, 0.1 * np.arange(9), y.real), atol=1e-6)


# -------------------------------------------------------------- samplers ----
def _abc(s2, s1, Em, En):
    a = abs(s2) ** 2 * abs(Em) ** 2 + abs(s1) ** 2 * abs(En) ** 2
    b = abs(s1) ** 2 * abs(Em) ** 2 + abs(s2) ** 2 * abs(En) ** 2
    c = 2.0 * (abs(s2) ** 2 - abs(s1) ** 2) * (Em * np.conj(En)).real
    return a, b, c


PHI_CASES = [  # (s2, s1, Em, En): isotropic, linear, elliptical with a strong sin*cos term
    (1.0 + 0.0j, 1.0 + 0.0j, 1.0 + 0.0j, 0.0j),
    (0.2 + 0.1j, 1.0 - 0.3j, 1.0 + 0.0j, 0.0j),
    (1.0 + 0.2j, 0.3 - 0.2j, math.cos(0.6) + 0j, math.sin(0.6) * np.exp(0.4j)),
    (0.05 + 0.0j, 0.9 + 0.4j, math.cos(2.2) + 0j, math.sin(2.2) * np.exp(-0.3j)),
]


@pytest.mark.parametrize("case", range(len(PHI_CASES)))
def test_rejection_azimuth_follows_F_and_its_trial_count(case):
    import azimuth_check as ac
    s2, s1, Em, En = PHI_CASES[case]
    a, b, c = _abc(s2, s1, Em, En)
    n = 200000
    phi, trials = oracle.phi_rejection(s2, s1, Em, En, seed=1000 + case, n=n)
    tol = 2.5 / math.sqrt(n)
    assert ac.cdf_discrepancy((a, b, c), np.cos(phi), np.sin(phi)) < tol
    # acceptance probability mean(F) / Fmax with mean(F) = (a + b) / 2
    fmax = 0.5 * (a + b) + 0.5 * math.hypot(a - b, c)
    p = 0.5 * (a + b) / fmax
    assert abs(n / trials - p) <= 5 * math.sqrt(p * (1 - p) / trials) + 1e-12
    if abs(c) > 0.2 * (a + b):
        with oracle.planted("phi_c_sign"):
            phi, _ = oracle.phi_rejection(s2, s1, Em, En, seed=1000 + case, n=n)
            assert ac.cdf_discrepancy((a, b, c), np.cos(phi), np.sin(phi)) > 5 * tol


@pytest.mark.parametrize("g", [0.9, 0.5, -0.3, 0.02])
def test_hg_inversion_is_exact(g):
    u = (np.arange(100000) + 0.5) / 100000
    mu = np.cos(oracle.sample_theta(_lib.PHASE_HG, g, u))
    cdf = (1 - g * g) / (2 * g) * (1 / np.sqrt(1 + g * g - 2 * g * mu) - 1 / (1 + g))
    assert np.allclose(cdf, u, rtol=0, atol=1e-9)
    assert abs(mu.mean() - g) < 1e-6
    with oracle.planted("hg_frac"):
        mu = np.cos(oracle.sample_theta(_lib.PHASE_HG, g, u))
        assert not np.allclose(np.nan_to_num(mu.mean(), nan=9.0), g, atol=1e-3)


def test_isotropic_theta_is_uniform_in_cosine():
    u = (np.arange(4096) + 0.5) / 4096
    mu = np.cos(oracle.sample_theta(_lib.PHASE_UNIFORM, 0.0, u))
    assert np.allclose((mu + 1) / 2, u, atol=1e-12)
    with oracle.planted("iso_theta"):
        mu = np.cos(oracle.sample_theta(_lib.PHASE_UNIFORM, 0.0, u))
        assert not np.allclose((mu + 1) / 2, u, atol=1e-3)


# ------------------------------------------------------------ scattering ----
def _random_frame(rng):
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 2] *= -1
    return q.T  # rows m, n, s; m x n = s


def _scatter_violations(P, E, th, ph, s2, s1):
    """Largest deviation of one step from its geometric definition."""
    P2, E2 = oracle.scatter_step(P, E, th, ph, s2, s1)
    m, n, s = P
    m2, n2, s2v = P2
    errs = [
        np.abs(P2 @ P2.T - np.eye(3)).max(), abs(np.linalg.det(P2) - 1.0), np.abs(np.cross(m2, n2) - s2v).max(),
        # the new direction's spherical angles in the old frame
        abs(s2v @ s - math.cos(th)), abs(s2v @ m - math.sin(th) * math.cos(ph)),
        abs(s2v @ n - math.sin(th) * math.sin(ph)),
        # n' is normal to the scattering plane at azimuth phi; m' lies in it
        np.abs(n2 - (-math.sin(ph) * m + math.cos(ph) * n)).max(), np.abs(m2 - np.cross(n2, s2v)).max(),
    ]
    # Jones: the field's parallel (n' x s) and perpendicular (n') components
    # scale by S2 and S1, then the vector is normalised (BH amplitude matrix)
    e = E[0] * m + E[1] * n
    e_par, e_perp = e @ np.cross(n2, s), e @ n2
    F = abs(s2 * e_par) ** 2 + abs(s1 * e_perp) ** 2
    errs += [abs(E2[0] - s2 * e_par / math.sqrt(F)), abs(E2[1] - s1 * e_perp / math.sqrt(F)),
             abs(np.linalg.norm(E2) - 1.0)]
    return max(errs)


def test_scatter_step_geometry_and_amplitudes():
    rng = np.random.default_rng(21)
    cases = []
    for _ in range(200):
        E = rng.normal(size=2) + 1j * rng.normal(size=2)
        cases.append((_random_frame(rng), E / np.linalg.norm(E), rng.uniform(0, math.pi), rng.uniform(0, 2 * math.pi),
                      complex(*rng.normal(size=2)), complex(*rng.normal(size=2))))
    assert max(_scatter_violations(*c) for c in cases) < 1e-12
    for plant in ("frame_transpose", "jones_swap"):
        with oracle.planted(plant):
            assert max(_scatter_violations(*c) for c in cases) > 1e-3, plant


# ---------------------------------------------------- transport by physics ----
class _IsoMedium(lmc.ScatteringMedium):
    """Isotropic scatterer with constant amplitudes S1 = S2 = 1 (unpolarised
    transport: the azimuth density is uniform and no weight depends on it)."""

    def __init__(self, mu_s: float, mu_a: float, mfp: float | None = None):
        super().__init__(lmc.UniformPhaseFunction())
        self.set_scattering_coefficient(mu_s)
        self.set_absorption_coefficient(mu_a)
        self.set_mean_free_path(mfp if mfp is not None else 1.0 / (mu_s + mu_a))

    def s_tables(self):
        return np.ones(2, complex), np.ones(2, complex), math.pi


def chandrasekhar_H(a: float, mu, n: int = 200):
    """H(mu) of isotropic scattering with albedo a, from
    1/H(mu) = sqrt(1 - a) + (a/2) int_0^1 x H(x) / (mu + x) dx by iteration on
    a Gauss-Legendre grid (Chandrasekhar, Radiative Transfer, ch. V)."""
    x, w = np.polynomial.legendre.leggauss(n)
    x, w = 0.5 * (x + 1.0), 0.5 * w
    H = np.ones(n)
    for _ in range(2000):
        Hn = 1.0 / (math.sqrt(1 - a) + 0.5 * a * ((x * w * H)[None, :] / (x[:, None] + x[None, :])).sum(axis=1))
        done = np.abs(Hn - H).max() < 1e-15
        H = Hn
        if done:
            break
    mu = np.atleast_1d(mu)
    return 1.0 / (math.sqrt(1 - a) + 0.5 * a * ((x * w * H)[None, :] / (mu[:, None] + x[None, :])).sum(axis=1)), (x, w, H)


def test_chandrasekhar_h_moment_identity():
    """Zeroth moment of H: int_0^1 H = (2/a)(1 - sqrt(1 - a))."""
    for a in (0.3, 0.9, 0.99):
        _, (x, w, H) = chandrasekhar_H(a, 1.0)
        assert abs((w * H).sum() - 2.0 / a * (1 - math.sqrt(1 - a))) < 1e-10


def _half_space(medium, sensors, n=1.0):
    sample = lmc.Sample(n)
    sample.add_layer(medium, 0.0, math.inf)
    laser = lmc.Laser(1.0, 0.0, WL, 0.0, lmc.LaserSource.Point)
    g = lmc.SensorsGroup()
    for s in sensors:
        g.add_detector(s)
    b = SceneBuild(sample, laser, g.detectors)
    b._keep = (medium, sample, laser, g)
    return b


@pytest.mark.parametrize("a", [0.9, 0.99])
def test_isotropic_half_space_reflectance_is_chandrasekhar(a):
    b = _half_space(_IsoMedium(a, 1 - a), [lmc.ReflectanceSensor(0.0, absorb=True)])
    n = 200000
    out, hits, ev, _ = oracle.run(b, n, seed=77, max_events=5000, n_threads=8)
    R = out[0] / n
    R_exact = 1.0 - chandrasekhar_H(a, 1.0)[0][0] * math.sqrt(1 - a)
    assert abs(R - R_exact) < 4.5 * 0.5 / math.sqrt(n), (R, R_exact)
    with oracle.planted("iso_theta"):
        out, _, _, _ = oracle.run(b, n // 4, seed=77, max_events=5000, n_threads=8)
        assert abs(out[0] / (n // 4) - R_exact) > 0.05


def _ballistic_sensor(z):
    s = lmc.ReflectanceSensor(z, absorb=False)
    s.set_direction_limit(lmc.CrossingDirection.Forward)
    s.set_events_limit(0, 0)
    return s


def test_ballistic_transmission_follows_the_mean_free_path_field():
    """Unscattered photons cross depth z with probability exp(-z / mfp), the
    layer's mean-free-path field (F6: here deliberately not 1/mu_t)."""
    zs = (0.3, 0.6, 1.2)
    b = _half_space(_IsoMedium(4.0, 1.0, mfp=0.5), [_ballistic_sensor(z) for z in zs])
    n = 100000
    out, hits, _, _ = oracle.run(b, n, seed=5, n_threads=8)
    for k, z in enumerate(zs):
        T = math.exp(-(z - 1e-9) / 0.5)
        assert abs(out[k] / n - T) < 4.5 * math.sqrt(T * (1 - T) / n), z
        assert hits[k] == out[k]  # weight 1: unscattered


def test_ballistic_transmission_through_two_layers_carries_optical_depth():
    """Layer 1 [0, 0.5): mu_t 2, mfp 0.4; layer 2 [0.5, inf): mu_t 6.  A step
    crossing the interface keeps its remaining optical depth (step - to
    interface) * mu_t1, converted at mu_t2: T(z) = exp(-0.5/mfp1) *
    exp(-(z - 0.5) mu_t2 / (mfp1 mu_t1)) (memoryless excess beyond 0.5)."""
    m1, m2 = _IsoMedium(1.8, 0.2, mfp=0.4), _IsoMedium(5.0, 1.0)
    sample = lmc.Sample(1.0)
    sample.add_layer(m1, 0.0, 0.5)
    sample.add_layer(m2, 0.5, math.inf)
    laser = lmc.Laser(1.0, 0.0, WL, 0.0, lmc.LaserSource.Point)
    g = lmc.SensorsGroup()
    g.add_detector(_ballistic_sensor(0.7))
    b = SceneBuild(sample, laser, g.detectors)
    n = 200000
    T = math.exp(-(0.5 - 1e-9) / 0.4) * math.exp(-0.2 * 6.0 / (0.4 * 2.0))
    out, _, _, _ = oracle.run(b, n, seed=9, n_threads=8)
    assert abs(out[0] / n - T) < 4.5 * math.sqrt(T * (1 - T) / n)
    with oracle.planted("rescale_old_mu"):
        out, _, _, _ = oracle.run(b, n, seed=9, n_threads=8)
        assert abs(out[0] / n - T) > 10 * math.sqrt(T * (1 - T) / n)


def test_identical_layers_equal_one_layer():
    """The same medium split into three layers: the interface sub-steps
    (rescale by the same mu_t, 1e-12 nudges) leave every trajectory as it was:
    same event, trial and hit counts, grids equal to rounding."""
    m, ph = scenes.mie_medium(0.5e-3, 1.59, 1.33, 10.0, 0.05, n_div=2000)

    def build(splits):
        sample = lmc.Sample(1.33)
        z = [0.0] + list(splits) + [math.inf]
        for lo, hi in zip(z[:-1], z[1:]):
            sample.add_layer(m, lo, hi)
        laser = lmc.Laser(1 / math.sqrt(2), 1j / math.sqrt(2), WL, 0.05, lmc.LaserSource.Gaussian)
        g = lmc.SensorsGroup()
        g.add_detector(lmc.FarFieldCBSSensor(0.2, 2 * math.pi, 0.0, 0.2 / 20, 2 * math.pi / 4, 0.0))
        g.add_detector(lmc.ReflectanceSensor(0.0, absorb=False))
        return SceneBuild(sample, laser, g.detectors)

    one, three = build([]), build([0.03, 0.11])
    a = oracle.run(one, 20000, seed=31, n_threads=4)
    b = oracle.run(three, 20000, seed=31, n_threads=4)
    assert a[2] == b[2] and a[3] == b[3] and np.array_equal(a[1], b[1])
    assert np.allclose(a[0], b[0], rtol=1e-9, atol=1e-12 * np.abs(a[0]).max())


# ---------------------------------------------------------------- sources ----
def _source_build(source, sigma, profile=lmc.TemporalProfile.Delta, pulse=0.0, rep=0.0, t0=0.0, pos=(0.0, 0.0)):
    m = _IsoMedium(1.0, 0.0)
    sample = lmc.Sample(1.0)
    sample.add_layer(m, 0.0, math.inf)
    laser = lmc.Laser(1.0, 0.0, WL, sigma, source)
    laser.position = lmc.Vec3(pos[0], pos[1], 1e-9)
    laser.set_temporal_profile(profile, pulse, rep, t0)
    b = SceneBuild(sample, laser, [])
    b._keep = (m, sample, laser)
    return b


def test_source_spatial_moments():
    n = 100000
    e = oracle.emit(_source_build(lmc.LaserSource.Gaussian, 0.2, pos=(0.3, -0.1)), 11, n)
    for k, c in ((0, 0.3), (1, -0.1)):
        assert abs(e[:, k].mean() - c) < 5 * 0.2 / math.sqrt(n)
        assert abs(e[:, k].var() - 0.04) < 5 * 0.04 * math.sqrt(2 / n)
    assert abs(np.corrcoef(e[:, 0], e[:, 1])[0, 1]) < 5 / math.sqrt(n)
    assert np.all(e[:, 2] == 1e-9)
    e = oracle.emit(_source_build(lmc.LaserSource.Uniform, 0.5), 12, n)
    r2 = e[:, 0] ** 2 + e[:, 1] ** 2
    assert r2.max() <= 0.25 and abs(r2.mean() - 0.125) < 5 * 0.25 * math.sqrt(1 / 12 / n)
    ang = np.mod(np.arctan2(e[:, 1], e[:, 0]), 2 * math.pi)
    assert abs(ang.mean() - math.pi) < 5 * math.pi / math.sqrt(3 * n)
    with oracle.planted("disk_no_sqrt"):
        e = oracle.emit(_source_build(lmc.LaserSource.Uniform, 0.5), 12, n)
        assert abs((e[:, 0] ** 2 + e[:, 1] ** 2).mean() - 0.125) > 0.02


@pytest.mark.parametrize("profile,pulse,rep,t0,mean,var", [
    (lmc.TemporalProfile.Gaussian, 50.0, 0.0, 10.0, 0.010, (0.050 / (2 * math.sqrt(2 * math.log(2)))) ** 2),
    (lmc.TemporalProfile.TopHat, 40.0, 0.0, 5.0, 0.005, 0.040 ** 2 / 12),
    (lmc.TemporalProfile.Exponential, 30.0, 0.0, 2.0, 0.032, 0.030 ** 2),
    (lmc.TemporalProfile.PulseTrain, 0.0, 80e6, 1.0, 0.001 + 0.5 * 1e9 / 80e6, (1e9 / 80e6) ** 2 / 12),
    (lmc.TemporalProfile.CW, 20.0, 0.0, 0.0, 0.010, 0.020 ** 2 / 12),
])
def test_emission_time_moments(profile, pulse, rep, t0, mean, var):
    """Launch times in ns from picosecond pulse parameters (laser.cpp:42-75)."""
    n = 100000
    t = oracle.emit(_source_build(lmc.LaserSource.Point, 0.0, profile, pulse, rep, t0), 13, n)[:, 3]
    assert abs(t.mean() - mean) < 5 * math.sqrt(var / n)
    assert abs(t.var() - var) < 6 * var * math.sqrt(2.5 / n)
    if profile == lmc.TemporalProfile.Gaussian:
        with oracle.planted("pulse_fwhm_as_sigma"):
            t = oracle.emit(_source_build(lmc.LaserSource.Point, 0.0, profile, pulse, rep, t0), 13, n)[:, 3]
            assert abs(t.var() - var) > 0.5 * var


# ---------------------------------------------------------------- sensors ----
def _fluence_scene(pos=(1.5, -0.7)):
    m, ph = scenes.mie_medium(0.5e-3, 1.59, 1.33, 10.0, 0.5, n_div=2000)
    sample = lmc.Sample(1.33)
    sample.add_layer(m, 0.0, math.inf)
    laser = lmc.Laser(math.cos(0.4), math.sin(0.4) * np.exp(0.7j), WL, 0.05, lmc.LaserSource.Gaussian)
    laser.position = lmc.Vec3(pos[0], pos[1], 1e-9)
    g = lmc.SensorsGroup()
    g.add_detector(lmc.PlanarFluenceSensor(0.0, 40.0, 40.0, 0.0, 40.0 / 160, 40.0 / 160, 0.0, absorb=False))
    g.add_detector(lmc.PhotonRecordSensor(0.0, absorb=True))
    b = SceneBuild(sample, laser, g.detectors)
    b._keep = (m, ph, sample, laser, g)
    return b


def _fluence_checks(b, n):
    out, hits, _, _ = oracle.run(b, n, seed=44, n_threads=8)
    rec = oracle.records()
    off, ln = b.offsets[0]
    S = out[off:off + ln].reshape(4, 160, 160)  # [S][x][y], one time row
    xc = -20.0 + (np.arange(160) + 0.5) * 0.25
    tot = S[0].sum()
    cx, cy = (S[0].sum(axis=1) * xc).sum() / tot, (S[0].sum(axis=0) * xc).sum() / tot
    # Stokes of the lab field (-Ex, Ey) from the records' Jones vectors and frames
    w = np.sqrt(rec[:, 6])
    Em, En = (rec[:, 26] + 1j * rec[:, 27]) * w, (rec[:, 28] + 1j * rec[:, 29]) * w
    Ex = -(Em * rec[:, 20] + En * rec[:, 23])
    Ey = Em * rec[:, 21] + En * rec[:, 24]
    inside = (np.abs(rec[:, 14]) < 20.0) & (np.abs(rec[:, 15]) < 20.0)
    st = np.array([abs(Ex) ** 2 + abs(Ey) ** 2, abs(Ex) ** 2 - abs(Ey) ** 2, 2 * (Ex * np.conj(Ey)).real,
                   2 * (Ex * np.conj(Ey)).imag])[:, inside].sum(axis=1)
    return cx, cy, S.reshape(4, -1).sum(axis=1), st


def test_planar_fluence_spot_and_lab_stokes():
    """Off-axis source at (1.5, -0.7) mm: the reflected fluence is centred on
    it (no flipped or swapped axes in the binning); the grid's Stokes totals
    are those of the lab field (-Ex, Ey) built from the PhotonRecordSensor's
    Jones vectors and frames of the same photons (detector.cpp:463-485, 812)."""
    b = _fluence_scene()
    cx, cy, grid, st = _fluence_checks(b, 40000)
    assert abs(cx - 1.5) < 0.03 and abs(cy + 0.7) < 0.03, (cx, cy)
    assert np.allclose(grid, st, rtol=1e-9, atol=1e-9 * abs(st[0]))
    assert abs(st[2]) > 1e-3 * st[0] and abs(st[3]) > 1e-3 * st[0]  # the x sign is visible in S2, S3
    with oracle.planted("fluence_pos_neg"):
        cx, cy, _, _ = _fluence_checks(b, 40000)
        assert abs(cx + 1.5) < 0.1
    with oracle.planted("fluence_no_xneg"):
        _, _, grid, st = _fluence_checks(b, 40000)
        assert not np.allclose(grid, st, rtol=1e-6)


def test_planar_field_phase_is_k0_times_optical_path():
    """Single photons: the field deposit divided by the lab Jones field
    sqrt(w) (E . P) equals exp(i k0 n L) with L the geometric path at the hit
    (the record's opticalpath) and n the sample index (detector.cpp:226-240)."""
    m, ph = scenes.mie_medium(0.5e-3, 1.59, 1.33, 10.0, 0.5, n_div=2000)
    sample = lmc.Sample(1.33)
    sample.add_layer(m, 0.0, math.inf)
    laser = lmc.Laser(1.0, 0.3j, WL, 0.0, lmc.LaserSource.Point)
    g = lmc.SensorsGroup()
    g.add_detector(lmc.PlanarFieldSensor(0.0, 8.0, 8.0, 8.0 / 64, 8.0 / 64, absorb=False))
    g.add_detector(lmc.PhotonRecordSensor(0.0, absorb=True))
    b = SceneBuild(sample, laser, g.detectors)
    k0 = 2 * math.pi / WL

    def worst(n_ph):
        bad, used = 0.0, 0
        for i in range(n_ph):
            out, hits, _, _ = oracle.run(b, 1, seed=3, photon_offset=i, n_threads=1)
            rec = oracle.records()
            if len(rec) != 1 or hits[0] != 1:
                continue
            r = rec[0]
            f = out[:2 * 64 * 64 * 2].reshape(2, 64 * 64, 2)
            ex = f[0, :, 0] + 1j * f[0, :, 1]
            j = int(np.argmax(np.abs(ex)))
            Ex = (r[26] + 1j * r[27]) * r[20] + (r[28] + 1j * r[29]) * r[23]
            if abs(Ex) < 1e-3:
                continue
            ratio = ex[j] / (Ex * math.sqrt(r[6]))
            bad = max(bad, abs(ratio - np.exp(1j * k0 * 1.33 * r[5])))
            used += 1
        return bad, used

    bad, used = worst(60)
    assert used >= 20 and bad < 1e-8
    with oracle.planted("field_phase_geom"):
        bad, _ = worst(20)
        assert bad > 0.1


def test_postprocess_farfield_solid_angles_and_scaling():
    """sum of dOmega over the grid = N_phi dphi (1 - cos(N_theta dtheta)); each
    processed value = raw / (N dOmega) (results.cpp:61-103, utils.cpp:76-82)."""
    det = lmc.FarFieldCBSSensor(0.3, 2 * math.pi, 0.0, 0.3 / 30, 2 * math.pi / 8, 0.0)
    rng = np.random.default_rng(5)
    det.buffer[:] = rng.uniform(size=det.buffer.size)
    pr = lmc.postprocess_farfield_cbs(det, 1234)
    dom = np.asarray(pr.dOmega)
    assert dom.shape == (30, 8)
    assert abs(dom.sum() - 8 * (2 * math.pi / 8) * (1 - math.cos(0.3))) < 1e-14
    assert np.allclose(dom[:, 0], (2 * math.pi / 8) * (np.cos(np.arange(30) * 0.01) - np.cos(np.arange(1, 31) * 0.01)),
                       rtol=1e-13)
    g = det.grids()
    for q, name in enumerate(("S0", "S1", "S2", "S3")):
        assert np.allclose(np.asarray(getattr(pr.coherent[0], name)), g[q, 0] / (1234 * dom), rtol=1e-14)
        assert np.allclose(np.asarray(getattr(pr.incoherent[0], name)), g[4 + q, 0] / (1234 * dom), rtol=1e-14)
