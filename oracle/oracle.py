"""Python access to the oracle (TEST INFRASTRUCTURE ONLY).

Loads oracle/liblmc_oracle.so, the FP64 C++ restatement of the reference's
transport path, and (where it was built) oracle/_ref/libmiev0.so, the
reference's own MIEV0 Mie solver compiled from its sources.  Two more builds
of the same source: liblmc_oracle_fast.so (release flags, the CPU baseline's
timing build) and liblmc_oracle_plant.so (planted mistakes for the tests'
sensitivity checks, switched by `planted`).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liblmc_oracle.so")
VARIANTS = {"parity": ORACLE_SO, "fast": os.path.join(HERE, "liblmc_oracle_fast.so"),
            "plant": os.path.join(HERE, "liblmc_oracle_plant.so")}
MIEV0_SO = os.path.join(HERE, "_ref", "libmiev0.so")

# planted mistakes of the plant build (enum Plant in lmc_oracle.cpp)
PLANTS = ("none", "xo_rot", "table_t", "data_t", "phi_c_sign", "hg_frac", "rescale_old_mu", "frame_transpose",
          "jones_swap", "disk_no_sqrt", "pulse_fwhm_as_sigma", "fluence_pos_neg", "fluence_no_xneg",
          "field_phase_geom", "iso_theta")

_dp = ctypes.POINTER(ctypes.c_double)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_libs = {}
_ref = None


def build() -> None:
    """Compile the oracle (and the MIEV0 reference build where its sources exist)."""
    subprocess.check_call(["make", "-C", HERE, "--no-print-directory"])


def lib(variant: str = "parity"):
    if variant not in _libs:
        path = VARIANTS[variant]
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.oracle_run.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_int, _dp, _u64p, _u64p, _u64p]
        L.oracle_last_error.restype = ctypes.c_char_p
        L.oracle_mie.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp, _dp, _dp,
                                 _dp, _dp, _dp]
        L.oracle_sampling_cdf.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.oracle_xoshiro_uniforms.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, _dp]
        L.oracle_xoshiro_u64.argtypes = [ctypes.c_uint64, ctypes.c_int, _u64p]
        L.oracle_mix_seed.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.oracle_species_i_norm.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.oracle_species_i_norm.restype = ctypes.c_double
        L.oracle_mix_seed.restype = ctypes.c_uint64
        L.oracle_reverse_field.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_double, _dp, _dp,
                                           _dp, _dp, _dp]
        L.oracle_records.argtypes = [_dp, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]
        L.oracle_philox4x32_10.argtypes = [ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32, ctypes.c_uint32,
                                           ctypes.POINTER(ctypes.c_uint32)]
        L.oracle_set_plant.argtypes = [ctypes.c_int]
        L.oracle_xoshiro_state_seq.argtypes = [_u64p, ctypes.c_int, _u64p, _u64p]
        L.oracle_table_sample.argtypes = [_dp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp,
                                          _dp]
        L.oracle_data_sample.argtypes = [_dp, ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp, _dp]
        L.oracle_phi_rejection.argtypes = [_dp, ctypes.c_uint64, ctypes.c_int, _dp, _u64p]
        L.oracle_sample_theta.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp, _dp]
        L.oracle_scatter_step.argtypes = [_dp, _dp, ctypes.c_double, ctypes.c_double, _dp]
        L.oracle_emit.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int64, _dp]
        _libs[variant] = L
    return _libs[variant]


_active = "parity"


class planted:
    """Context manager: route this module's calls to the plant build with
    mistake `name` (one of PLANTS) switched on."""

    def __init__(self, name: str):
        self.k = PLANTS.index(name)

    def __enter__(self):
        global _active
        L = lib("plant")
        if L.oracle_set_plant(self.k) != 1:
            raise RuntimeError("liblmc_oracle_plant.so was built without LMC_ORACLE_PLANTS")
        _active = "plant"
        return self

    def __exit__(self, *exc):
        global _active
        lib("plant").oracle_set_plant(0)
        _active = "parity"
        return False


def _cur():
    return lib(_active)


def run(scene_build, n_photons: int, seed: int, max_events: int = 1000, track: bool = True,
        n_threads: int = 0, photon_offset: int = 0, variant: str | None = None):
    """Run the restatement on a flattened scene (luminis_mc.simulation.SceneBuild).
    Returns (accumulators, hits, events, phi_trials).  variant: None = the
    parity build (or the plant build inside `planted`), "fast" = the release
    build the CPU baseline times."""
    L = lib(variant) if variant else _cur()
    out = np.zeros(max(1, scene_build.n_out))
    hits = np.zeros(max(1, len(scene_build.sensors)), dtype=np.uint64)
    ev = ctypes.c_uint64(0)
    tr = ctypes.c_uint64(0)
    rc = L.oracle_run(ctypes.addressof(scene_build.desc), int(n_photons), int(photon_offset),
                      int(seed) & ((1 << 64) - 1), int(max_events), int(bool(track)), int(n_threads),
                      out.ctypes.data_as(_dp), hits.ctypes.data_as(_u64p), ctypes.byref(ev), ctypes.byref(tr))
    if rc != 0:
        raise ValueError(L.oracle_last_error().decode())
    return out[:scene_build.n_out], hits, ev.value, tr.value


def records() -> np.ndarray:
    """PhotonRecordSensor rows of the last run() ((n, 34), column 0 the
    sensor's index; layout of LMC_RECORD_FIELDS)."""
    L = _cur()
    n = ctypes.c_int64(0)
    L.oracle_records(None, 0, ctypes.byref(n))
    rows = np.zeros((n.value, 34))
    if n.value:
        L.oracle_records(rows.ctypes.data_as(_dp), n.value, ctypes.byref(n))
    return rows


def mie(x: float, m: complex, mu):
    mu = np.ascontiguousarray(np.atleast_1d(mu), dtype=np.float64)
    s1 = np.zeros(2 * mu.size)
    s2 = np.zeros(2 * mu.size)
    qe, qs, gq = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    _cur().oracle_mie(float(x), float(m.real), float(m.imag), mu.size, mu.ctypes.data_as(_dp),
                     s1.ctypes.data_as(_dp), s2.ctypes.data_as(_dp), ctypes.byref(qe), ctypes.byref(qs),
                     ctypes.byref(gq))
    return s1[0::2] + 1j * s1[1::2], s2[0::2] + 1j * s2[1::2], qe.value, qs.value, gq.value


def sampling_cdf(n_div: int, min_val: float, max_val: float, pdf) -> np.ndarray:
    pdf = np.ascontiguousarray(pdf, dtype=np.float64)
    cdf = np.zeros(n_div + 1)
    rc = _cur().oracle_sampling_cdf(int(n_div), float(min_val), float(max_val), pdf.ctypes.data_as(_dp),
                                   cdf.ctypes.data_as(_dp))
    if rc != 0:
        raise ValueError("oracle_sampling_cdf failed")
    return cdf


def xoshiro_uniforms(seed: int, idx: int, n: int) -> np.ndarray:
    out = np.zeros(n)
    _cur().oracle_xoshiro_uniforms(int(seed), int(idx), int(n), out.ctypes.data_as(_dp))
    return out


def xoshiro_u64(state_seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.uint64)
    _cur().oracle_xoshiro_u64(int(state_seed), int(n), out.ctypes.data_as(_u64p))
    return out


def mix_seed(base: int, tid: int) -> int:
    return int(_cur().oracle_mix_seed(int(base), int(tid)))


def reverse_field(P0, P1, Pnm1, Pn, T, E0, s_last, s_first, dtheta: float):
    """The restatement's reverse_field on explicit inputs.  P*: 3x3 frames (rows
    m, n, s); T: 2x2 complex raw interior product; E0: launch Jones vector;
    s_last / s_first: (S1, S2) complex tables on theta_i = i*dtheta of the last-
    and first-vertex species.  Returns the normalised reverse field (2 complex)
    in the exit basis."""
    def c(a):
        a = np.asarray(a, dtype=np.complex128).ravel()
        return np.ascontiguousarray(np.stack([a.real, a.imag], axis=1).ravel())

    arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (P0, P1, Pnm1, Pn)]
    Tc, Ec = c(T), c(E0)
    s1l, s2l, s1f, s2f = c(s_last[0]), c(s_last[1]), c(s_first[0]), c(s_first[1])
    out = np.zeros(4)
    _cur().oracle_reverse_field(*(a.ctypes.data_as(_dp) for a in arrs), Tc.ctypes.data_as(_dp), Ec.ctypes.data_as(_dp),
                               int(len(s_last[0])), float(dtheta), s1l.ctypes.data_as(_dp), s2l.ctypes.data_as(_dp),
                               s1f.ctypes.data_as(_dp), s2f.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
    return np.array([out[0] + 1j * out[1], out[2] + 1j * out[3]])


def species_i_norm(scene_build, sp: int) -> float:
    """compute_I_norm (detector.cpp:105-118) of species `sp` of a scene build."""
    return float(_cur().oracle_species_i_norm(ctypes.addressof(scene_build.desc), int(sp)))


def philox(ctr, k0: int, k1: int):
    c = (ctypes.c_uint32 * 4)(*[int(v) & 0xFFFFFFFF for v in ctr])
    o = (ctypes.c_uint32 * 4)()
    _cur().oracle_philox4x32_10(c, int(k0) & 0xFFFFFFFF, int(k1) & 0xFFFFFFFF, o)
    return [int(v) for v in o]


def xoshiro_state_seq(state, n: int):
    """xoshiro256++ from the raw state (4 words): (outputs[n], states[n, 4])."""
    s = np.ascontiguousarray(np.asarray(state, dtype=np.uint64))
    out = np.zeros(n, dtype=np.uint64)
    st = np.zeros((n, 4), dtype=np.uint64)
    _cur().oracle_xoshiro_state_seq(s.ctypes.data_as(_u64p), int(n), out.ctypes.data_as(_u64p),
                                    st.ctypes.data_as(_u64p))
    return out, st


def table_sample(cdf, min_val: float, step: float, u):
    c = np.ascontiguousarray(cdf, dtype=np.float64)
    u = np.ascontiguousarray(np.atleast_1d(u), dtype=np.float64)
    out = np.zeros(u.size)
    _cur().oracle_table_sample(c.ctypes.data_as(_dp), c.size, float(min_val), float(step), u.size,
                               u.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
    return out


def data_sample(y, dx: float, x):
    y = np.asarray(y, dtype=np.complex128)
    yy = np.ascontiguousarray(np.stack([y.real, y.imag], axis=1).ravel())
    x = np.ascontiguousarray(np.atleast_1d(x), dtype=np.float64)
    out = np.zeros(2 * x.size)
    _cur().oracle_data_sample(yy.ctypes.data_as(_dp), y.size, float(dx), x.size, x.ctypes.data_as(_dp),
                              out.ctypes.data_as(_dp))
    return out[0::2] + 1j * out[1::2]


def phi_rejection(s2: complex, s1: complex, Em: complex, En: complex, seed: int, n: int):
    """n azimuths of the rejection sampler; returns (phi, trials)."""
    amp = np.array([s2.real, s2.imag, s1.real, s1.imag, Em.real, Em.imag, En.real, En.imag])
    phi = np.zeros(n)
    tr = ctypes.c_uint64(0)
    _cur().oracle_phi_rejection(amp.ctypes.data_as(_dp), int(seed), int(n), phi.ctypes.data_as(_dp),
                                ctypes.byref(tr))
    return phi, tr.value


def sample_theta(kind: int, g: float, u):
    u = np.ascontiguousarray(np.atleast_1d(u), dtype=np.float64)
    out = np.zeros(u.size)
    _cur().oracle_sample_theta(int(kind), float(g), u.size, u.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
    return out


def scatter_step(P, E, theta: float, phi: float, s2: complex, s1: complex):
    """One frame + Jones scattering step; returns (P', E')."""
    Pm = np.ascontiguousarray(np.asarray(P, dtype=np.float64).reshape(9).copy())
    Ev = np.asarray(E, dtype=np.complex128)
    Ea = np.ascontiguousarray([Ev[0].real, Ev[0].imag, Ev[1].real, Ev[1].imag])
    amp = np.array([s2.real, s2.imag, s1.real, s1.imag])
    _cur().oracle_scatter_step(Pm.ctypes.data_as(_dp), Ea.ctypes.data_as(_dp), float(theta), float(phi),
                               amp.ctypes.data_as(_dp))
    return Pm.reshape(3, 3), np.array([Ea[0] + 1j * Ea[1], Ea[2] + 1j * Ea[3]])


def emit(scene_build, seed: int, n: int, offset: int = 0) -> np.ndarray:
    """Launch state (x, y, z, launch time) of photons [offset, offset + n)."""
    out = np.zeros((n, 4))
    rc = _cur().oracle_emit(ctypes.addressof(scene_build.desc), int(seed), int(offset), int(n),
                            out.ctypes.data_as(_dp))
    if rc != 0:
        raise ValueError(_cur().oracle_last_error().decode())
    return out


# ---------------------------------------------------------------- MIEV0 --
class _C(ctypes.Structure):
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]


def ref_available() -> bool:
    return os.path.exists(MIEV0_SO)


def ref():
    global _ref
    if _ref is None:
        R = ctypes.CDLL(MIEV0_SO)
        R.miev.argtypes = [ctypes.c_double, _C, ctypes.c_int, _dp, _dp, _dp]
        R.mievinfo.argtypes = [ctypes.c_double, _C, _dp, _dp, _dp]
        _ref = R
    return _ref


def ref_miev(x: float, m: complex, mu):
    """The reference's miev() wrapper (src/mie/dmiev.cpp:49-62) over MIEV0."""
    mu = np.ascontiguousarray(np.atleast_1d(mu), dtype=np.float64)
    s1 = np.zeros(2 * mu.size)
    s2 = np.zeros(2 * mu.size)
    ref().miev(float(x), _C(m.real, m.imag), mu.size, mu.ctypes.data_as(_dp), s1.ctypes.data_as(_dp),
               s2.ctypes.data_as(_dp))
    return s1[0::2] + 1j * s1[1::2], s2[0::2] + 1j * s2[1::2]


def ref_mievinfo(x: float, m: complex):
    qe, qs, g = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    ref().mievinfo(float(x), _C(m.real, m.imag), ctypes.byref(qe), ctypes.byref(qs), ctypes.byref(g))
    return qe.value, qs.value, g.value
