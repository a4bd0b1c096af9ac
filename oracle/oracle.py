"""Python access to the oracle (TEST INFRASTRUCTURE ONLY).

Loads oracle/liblmc_oracle.so, the FP64 C++ restatement of the reference's
transport path, and (where it was built) oracle/_ref/libmiev0.so, the
reference's own MIEV0 Mie solver compiled from its sources.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liblmc_oracle.so")
MIEV0_SO = os.path.join(HERE, "_ref", "libmiev0.so")

_dp = ctypes.POINTER(ctypes.c_double)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_lib = None
_ref = None


def build() -> None:
    """Compile the oracle (and the MIEV0 reference build where its sources exist)."""
    subprocess.check_call(["make", "-C", HERE, "--no-print-directory"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = ctypes.CDLL(ORACLE_SO)
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
        _lib = L
    return _lib


def run(scene_build, n_photons: int, seed: int, max_events: int = 1000, track: bool = True,
        n_threads: int = 0, photon_offset: int = 0):
    """Run the restatement on a flattened scene (luminis_mc.simulation.SceneBuild).
    Returns (accumulators, hits, events, phi_trials)."""
    L = lib()
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
    L = lib()
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
    lib().oracle_mie(float(x), float(m.real), float(m.imag), mu.size, mu.ctypes.data_as(_dp),
                     s1.ctypes.data_as(_dp), s2.ctypes.data_as(_dp), ctypes.byref(qe), ctypes.byref(qs),
                     ctypes.byref(gq))
    return s1[0::2] + 1j * s1[1::2], s2[0::2] + 1j * s2[1::2], qe.value, qs.value, gq.value


def sampling_cdf(n_div: int, min_val: float, max_val: float, pdf) -> np.ndarray:
    pdf = np.ascontiguousarray(pdf, dtype=np.float64)
    cdf = np.zeros(n_div + 1)
    rc = lib().oracle_sampling_cdf(int(n_div), float(min_val), float(max_val), pdf.ctypes.data_as(_dp),
                                   cdf.ctypes.data_as(_dp))
    if rc != 0:
        raise ValueError("oracle_sampling_cdf failed")
    return cdf


def xoshiro_uniforms(seed: int, idx: int, n: int) -> np.ndarray:
    out = np.zeros(n)
    lib().oracle_xoshiro_uniforms(int(seed), int(idx), int(n), out.ctypes.data_as(_dp))
    return out


def xoshiro_u64(state_seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.uint64)
    lib().oracle_xoshiro_u64(int(state_seed), int(n), out.ctypes.data_as(_u64p))
    return out


def mix_seed(base: int, tid: int) -> int:
    return int(lib().oracle_mix_seed(int(base), int(tid)))


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
    lib().oracle_reverse_field(*(a.ctypes.data_as(_dp) for a in arrs), Tc.ctypes.data_as(_dp), Ec.ctypes.data_as(_dp),
                               int(len(s_last[0])), float(dtheta), s1l.ctypes.data_as(_dp), s2l.ctypes.data_as(_dp),
                               s1f.ctypes.data_as(_dp), s2f.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
    return np.array([out[0] + 1j * out[1], out[2] + 1j * out[3]])


def species_i_norm(scene_build, sp: int) -> float:
    """compute_I_norm (detector.cpp:105-118) of species `sp` of a scene build."""
    return float(lib().oracle_species_i_norm(ctypes.addressof(scene_build.desc), int(sp)))


def philox(ctr, k0: int, k1: int):
    c = (ctypes.c_uint32 * 4)(*[int(v) & 0xFFFFFFFF for v in ctr])
    o = (ctypes.c_uint32 * 4)()
    lib().oracle_philox4x32_10(c, int(k0) & 0xFFFFFFFF, int(k1) & 0xFFFFFFFF, o)
    return [int(v) for v in o]


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
