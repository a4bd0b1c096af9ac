"""Distribution check for the azimuth sampler of the production path.

The density is F(phi) = a cos^2 phi + b sin^2 phi + c sin phi cos phi
(src/phase.cpp:27-49).  The sampler maps two words to phi; fed a regular
lattice of its two uniforms and both half-period bits, the lattice image must
reproduce F's CDF, whose bin masses follow from the antiderivative
a (p/2 + sin 2p / 4) + b (p/2 - sin 2p / 4) + c sin^2 p / 2.
"""
import numpy as np

N_LATTICE = 512
N_BINS = 128
CDF_TOL = 2e-3


def forms(rng, n):
    """(a, b, c) intensity forms: isotropic, fully polarised, random PSD, near-singular."""
    out = [(1.0, 1.0, 0.0), (1.0, 0.0, 0.0), (0.0, 2.0, 0.0), (1.0, 1.0, 2.0), (0.3, 0.7, -2 * np.sqrt(0.21)),
           (1.0, 0.99, 0.01)]
    while len(out) < n:
        a, b = rng.uniform(0.0, 2.0, size=2)
        c = rng.uniform(-1.0, 1.0) * 2.0 * np.sqrt(a * b) * (1.0 - 10 ** rng.uniform(-7, 0))
        out.append((a, b, c))
    return np.array(out)


def lattice_words(n=N_LATTICE):
    """Word pairs (component/radius word, angle word) on an n x n x 2 lattice."""
    k = np.arange(n, dtype=np.uint64)
    wv = ((2 * k + 1) << np.uint64(23 - int(np.log2(n)))) << np.uint64(8)   # v = (i + 1/2) / n
    wa = ((2 * k + 1) << np.uint64(22 - int(np.log2(n)))) << np.uint64(9)   # angle = (j + 1/2) / n turns
    V, A = np.meshgrid(wv, wa, indexing="ij")
    V, A = V.ravel(), A.ravel()
    words = np.stack([np.concatenate([V, V]), np.concatenate([A, A | np.uint64(0x100)])], axis=1)
    return words.astype(np.uint32)


def exact_bins(a, b, c, nb=N_BINS):
    e = np.linspace(0.0, 2.0 * np.pi, nb + 1)
    I = a * (e / 2 + np.sin(2 * e) / 4) + b * (e / 2 - np.sin(2 * e) / 4) + c * np.sin(e) ** 2 / 2
    w = np.diff(I)
    return w / w.sum()


def cdf_discrepancy(form, cph, sph, nb=N_BINS):
    """Largest |empirical CDF - exact CDF| over nb bin edges of [0, 2 pi)."""
    phi = np.mod(np.arctan2(sph, cph), 2.0 * np.pi)
    hist = np.histogram(phi, bins=nb, range=(0.0, 2.0 * np.pi))[0] / phi.size
    return float(np.abs(np.cumsum(hist) - np.cumsum(exact_bins(*form, nb))).max())
