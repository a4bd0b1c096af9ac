"""Gradients of detector outputs with respect to layer parameters.

The GPU transport (lmc_scene_run_adjoint) accumulates, beside every detector
value, its pathwise score derivative with respect to each layer's scattering
and absorption coefficients and, for Henyey-Greenstein layers, the asymmetry
g.  With the free path drawn as -ln(u) / mu_t (mean free path tracking
1/mu_t) and the weight multiplied by the albedo mu_s/mu_t at each scatter, a
path's density times weight is prod(mu_s of each scatter) * exp(-sum_l mu_t,l
L_l), so for every detected photon

    d ln / d mu_s,l = N_l / mu_s,l - L_l,     d ln / d mu_a,l = -L_l,
    d ln / d g_l    = sum over HG scatters in l of d ln p_HG(cos theta; g) / d g

(N_l scatters and L_l geometric path in layer l up to the detector).  The
gradient of the expectation is the expectation of value x score, estimated
from the same photons as the value itself.  The reference has no gradient
path; its MiePhaseFunction g is a derived quantity, not a parameter, so the g
row of a non-HG layer is zero.

`DifferentiableTransport` exposes this as a torch.autograd.Function over a
(n_layers, 3) tensor of (mu_s, mu_a, g) per layer; `crn_central_difference`
is the finite-difference check (common random numbers: the same per-photon
streams at theta +- h).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .medium import Sample
from .sensors import Laser
from .simulation import GpuScene, SceneBuild

PARAMS = ("mu_s", "mu_a", "g")


class _ParamBuild:
    """A SceneBuild with its layers' (mu_s, mu_a, g) replaced; tables shared."""

    def __init__(self, base: SceneBuild, theta: np.ndarray):
        nl = base.desc.n_layers
        self.tables = base.tables
        self.sensors = base.sensors
        self.sensor_descs = base.sensor_descs
        self.offsets = base.offsets
        self.n_out = base.n_out
        self.layers = (_lib.LayerDesc * nl)()
        ctypes.memmove(self.layers, base.layers, ctypes.sizeof(self.layers))
        # A species shared by several layers (one medium object added to
        # several layers) gets a copy per extra layer, so each layer's
        # (mu_s, mu_a, g) stays its own; the copies share the tables.
        ns = base.desc.n_species
        first = [base.layers[l].species_first for l in range(nl)]
        extra = [l for l in range(nl) if first.index(first[l]) != l]
        if ns + len(extra) > _lib.MAX_SPECIES:
            raise NotImplementedError(f"gradients: {ns + len(extra)} species after giving each layer its own "
                                      f"(at most {_lib.MAX_SPECIES})")
        self.species = (_lib.SpeciesDesc * (ns + len(extra)))()
        ctypes.memmove(self.species, base.species, ctypes.sizeof(_lib.SpeciesDesc) * ns)
        for k, l in enumerate(extra):
            ctypes.memmove(ctypes.byref(self.species[ns + k]), ctypes.byref(base.species[first[l]]),
                           ctypes.sizeof(_lib.SpeciesDesc))
            self.layers[l].species_first = ns + k
        for l in range(nl):
            mu_s, mu_a, g = (float(v) for v in theta[l])
            if not (mu_s > 0.0 and mu_a >= 0.0):
                raise ValueError(f"layer {l}: need mu_s > 0 and mu_a >= 0")
            L = self.layers[l]
            L.mu_s, L.mu_a, L.mu_t = mu_s, mu_a, mu_s + mu_a
            L.mfp = 1.0 / (mu_s + mu_a)
            S = self.species[L.species_first]
            S.mu_s, S.mu_a, S.mfp = mu_s, mu_a, 1.0 / (mu_s + mu_a)
            if S.phase_kind == _lib.PHASE_HG:
                if not (-1.0 < g < 1.0) or g == 0.0:
                    raise ValueError(f"layer {l}: Henyey-Greenstein g must lie in (-1, 1) and be non-zero")
                S.hg_g = g
        d = _lib.SceneDesc()
        ctypes.memmove(ctypes.byref(d), ctypes.byref(base.desc), ctypes.sizeof(d))
        d.layers = ctypes.cast(self.layers, ctypes.POINTER(_lib.LayerDesc))
        d.species = ctypes.cast(self.species, ctypes.POINTER(_lib.SpeciesDesc))
        d.n_species = ns + len(extra)
        self.desc = d


class DifferentiableTransport:
    """Detector outputs per launched photon as a differentiable function of
    the layers' (mu_s, mu_a, g).

    Call with a torch tensor of shape (n_layers, 3) to get a float64 tensor of
    the scene's flat accumulators (sensor after sensor, each in the layout of
    its grid) divided by n_photons; backward() applies the score gradients.
    Every call runs photons [photon_offset, photon_offset + n_photons) of the
    same per-photon random streams (common random numbers across parameter values)."""

    def __init__(self, sample: Sample, laser: Laser, sensors, n_photons: int, seed: int = 0,
                 max_events: int = 1000, track_reverse_paths: bool = True, precision: str = "fp32",
                 photon_offset: int = 0):
        self.base = SceneBuild(sample, laser, list(sensors))
        nl = self.base.desc.n_layers
        if nl > 4:
            raise NotImplementedError("gradients are tracked for at most 4 layers")
        for l in range(nl):
            if self.base.layers[l].species_count != 1:
                raise NotImplementedError("gradients need homogeneous layers (one species each)")
        self.n_layers = nl
        self.n_photons = int(n_photons)
        self.seed = int(seed)
        self.max_events = int(max_events)
        self.track = bool(track_reverse_paths)
        self.precision = precision
        self.photon_offset = int(photon_offset)
        self.hg_layers = [self.base.species[self.base.layers[l].species_first].phase_kind == _lib.PHASE_HG
                          for l in range(nl)]
        self.last_stats = None

    @property
    def n_out(self) -> int:
        return self.base.n_out

    def theta0(self) -> np.ndarray:
        """(n_layers, 3) parameters of the scene as built (g = 0 for non-HG layers)."""
        th = np.zeros((self.n_layers, 3))
        for l in range(self.n_layers):
            L = self.base.layers[l]
            th[l, 0], th[l, 1] = L.mu_s, L.mu_a
            S = self.base.species[L.species_first]
            th[l, 2] = S.hg_g if S.phase_kind == _lib.PHASE_HG else 0.0
        return th

    def _run_desc(self, n: int, offset: int) -> _lib.RunDesc:
        r = _lib.RunDesc()
        r.n_photons = int(n)
        r.photon_offset = int(offset)
        r.seed = self.seed & ((1 << 64) - 1)
        r.max_events = self.max_events
        r.track_reverse_paths = int(self.track)
        r.precision = _lib.PREC_FP32 if self.precision == "fp32" else _lib.PREC_FP64
        r.rng = _lib.RNG_PHILOX
        return r

    def scene(self, theta) -> GpuScene:
        """A device scene at parameters theta (reuse it for several runs)."""
        return GpuScene.from_build(_ParamBuild(self.base, np.asarray(theta, dtype=np.float64)))

    def evaluate(self, theta, with_grad: bool = True, n_photons: int | None = None,
                 photon_offset: int | None = None, scene: GpuScene | None = None):
        """(values, jacobian) per photon: values (n_out,), jacobian (n_layers, 3, n_out) or None."""
        n = self.n_photons if n_photons is None else int(n_photons)
        off = self.photon_offset if photon_offset is None else int(photon_offset)
        sc = scene if scene is not None else self.scene(theta)
        try:
            sc.zero()
            r = self._run_desc(n, off)
            if with_grad:
                grad, st = sc.run_adjoint(r)
                jac = grad.reshape(self.n_layers, 3, self.n_out) / n
            else:
                st = sc.run(r)
                jac = None
            out, _ = sc.read()
            self.last_stats = st
            return out / n, jac
        finally:
            if scene is None:
                sc.close()

    def __call__(self, theta):
        return _transport_fn().apply(theta, self)


_FN = None


def _transport_fn():
    """torch.autograd.Function wrapping DifferentiableTransport.evaluate (built on first use)."""
    global _FN
    if _FN is None:
        import torch

        class TransportFunction(torch.autograd.Function):
            @staticmethod
            def forward(ctx, theta, tr):
                # outputs and the Jacobian live on theta's device (CPU or GPU)
                th = theta.detach().to("cpu", torch.float64).numpy()
                out, jac = tr.evaluate(th, with_grad=theta.requires_grad)
                if jac is not None:
                    ctx.save_for_backward(torch.from_numpy(jac).to(theta.device))
                ctx.theta_meta = (theta.dtype, theta.device)
                return torch.from_numpy(out).to(theta.device)

            @staticmethod
            def backward(ctx, g_out):
                (jac,) = ctx.saved_tensors
                dtype, device = ctx.theta_meta
                g = torch.tensordot(jac, g_out.to(jac.device, torch.float64), dims=([2], [0]))
                return g.to(device=device, dtype=dtype), None

        _FN = TransportFunction
    return _FN


def crn_central_difference(tr: DifferentiableTransport, theta, layer: int, param: int, rel_step: float = 1e-3,
                           n_photons: int | None = None, photon_offset: int | None = None, scenes=None):
    """Central finite difference of the per-photon outputs with respect to
    theta[layer, param], step rel_step * |theta|, on the same photons (CRN).
    scenes: optional (scene_plus, scene_minus) already built at theta +- h."""
    th = np.array(theta, dtype=np.float64)
    h = rel_step * abs(th[layer, param])
    if h == 0.0:
        raise ValueError("parameter is zero: relative step undefined")
    tp, tm = th.copy(), th.copy()
    tp[layer, param] += h
    tm[layer, param] -= h
    sp, sm = scenes if scenes is not None else (None, None)
    fp, _ = tr.evaluate(tp, with_grad=False, n_photons=n_photons, photon_offset=photon_offset, scene=sp)
    fm, _ = tr.evaluate(tm, with_grad=False, n_photons=n_photons, photon_offset=photon_offset, scene=sm)
    return (fp - fm) / (2.0 * h)


def perturbed(theta, layer: int, param: int, rel_step: float = 1e-3):
    """(theta + h e, theta - h e) for the central difference above."""
    th = np.array(theta, dtype=np.float64)
    h = rel_step * abs(th[layer, param])
    tp, tm = th.copy(), th.copy()
    tp[layer, param] += h
    tm[layer, param] -= h
    return tp, tm


__all__ = ["DifferentiableTransport", "crn_central_difference", "perturbed", "PARAMS"]
