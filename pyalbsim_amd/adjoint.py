"""Waveform-loss gradients of the batched form by path replay.

Parameters per scene, theta = (a, b, g, depth, albedo, roughness):
absorption and scattering coefficients (1/m), Henyey-Greenstein asymmetry,
water depth (m), floor albedo and surface roughness r (GGX alpha = r^2), the
quantities the reference's SimulationConfig holds (config/water.py,
sea_floor.py, sea_surface.py) plus the HG asymmetry of the sweep.

Estimator.  The forward pass adds c = E_q E_p K mult for every (sensor-path
vertex q, k-nearest map record p) into bin(t_q + t_p).  With the sampled paths
held fixed (detached sampling; the k-NN selection and radius K are piecewise
constant in theta and held fixed too),

    d c / d theta = c (d log mult + S_p + S_q),

where S of a path prefix with n scatters already weighted by omega0, a water
length L whose optical depth was accumulated, m floor reflections and
phase-sampling score sum G is (alpha = alpha_L, the lidar attenuation used as
free-path rate; omega0 = (alpha - a) / alpha; [scatter] = 1 when the record is
itself a scatter vertex, whose free-path density alpha e^{-alpha L} counts):

    S = ([scatter] / alpha + n / (alpha - a) - L) d alpha
        - n / (alpha - a) d a |alpha  +  m / albedo d albedo  +  G d g.

alpha depends on (a, b, g) through the Churnside formula
(physics/models/water.py:163-182, restated in model.lidar_attenuation; g
enters through the HG backscatter fraction).  d log mult is the phase-mass
derivative (scatter map, g), the GGX BRDF derivative (surface reflection,
roughness) and 1 / albedo (bottom).

The GPU work is: one adjoint-gather launch per photon type (the adjoint
weights w = G[bin] c to records, queries and per-scene direct sums), one
radix sort per sink to index records by (photon, event), and one replay trace
per pass that recomputes every path from its Philox stream and reduces
w * S per scene.  No path is stored.

Depth moves the seafloor plane: the paths' support changes, which detached
sampling does not cover.  Its derivative is taken by a difference with
common random numbers: one-sided from the forward pass's own waveforms (one
more batched forward pass, depth_grad="fd1", the default), central (two
more, "fd"), or left out ("none").
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import abi, runtime
from .model import (SCATTER_DIVISIONS, henyey_greenstein_backscatter, henyey_greenstein_table, lidar_attenuation,
                    phase_bin_mass)
from .sweep import shot_config

PARAMS = ("a", "b", "g", "depth", "albedo", "roughness")
REPLAY_SUMS = 6  # per-scene sums of alb_replay_batched


def phase_g_tables(g: float, h: float = 1e-5):
    """(score, dlog_mass) of the HG table at g: d log of the inverse-CDF
    segment probabilities ([0] the atom cdf[0], [i] cdf[i] - cdf[i-1]) and of
    the gather's per-bin phase masses, by central differences in g."""
    def seg(gg):
        th, _, cdf = henyey_greenstein_table(gg)
        p = np.empty_like(cdf)
        p[0] = cdf[0]
        p[1:] = np.diff(cdf)
        return th, cdf, p

    thp, cdp, pp = seg(g + h)
    thm, cdm, pm = seg(g - h)
    with np.errstate(divide="ignore", invalid="ignore"):
        score = (np.log(pp) - np.log(pm)) / (2 * h)
        mp = phase_bin_mass(thp, cdp)
        mm = phase_bin_mass(thm, cdm)
        dlm = (np.log(mp) - np.log(mm)) / (2 * h)
    score[~np.isfinite(score)] = 0.0
    dlm[~np.isfinite(dlm)] = 0.0
    assert len(score) == SCATTER_DIVISIONS
    return score, dlm


@dataclass
class SceneDerivs:
    alpha: float     # float32 lidar attenuation used by the kernels
    a: float
    albedo: float
    roughness: float
    dalpha: np.ndarray  # d alpha / d (a, b, g)


def scene_derivatives(model, theta) -> SceneDerivs:
    """d alpha_L / d(a, b, g) by central differences of the Churnside formula."""
    a, b, g, depth, albedo, rough = [float(v) for v in theta]
    layer = model.config.water.layers[0]
    H = model.config.scene.flying_height
    lam = model.config.laser.wavelength
    div = model.config.laser.divergence_angle

    def al(aa, bb, gg):
        return lidar_attenuation(aa, bb, layer.salinity, lam, H, div, henyey_greenstein_backscatter(gg))

    d = np.zeros(3)
    for i, (x, f) in enumerate(((a, lambda v: al(v, b, g)), (b, lambda v: al(a, v, g)),
                                (g, lambda v: al(a, b, v)))):
        h = 1e-6 * max(abs(x), 1e-3)
        d[i] = (f(x + h) - f(x - h)) / (2 * h)
    return SceneDerivs(model.alpha_l, a, albedo, rough, d)


def model_from_theta(theta, sample_rate=None, tilt=None):
    """SceneModel of the sweep geometry for theta = (a, b, g, depth, albedo,
    roughness); sample_rate: receiver sample rate (default 1e10 /s); tilt:
    optional (slope, azimuth) of the floor in degrees."""
    from .model import build_scene

    a, b, g, depth, albedo, rough = [float(v) for v in theta]
    sr = int(1e10) if sample_rate is None else int(sample_rate)
    sl, az = (0.0, 0.0) if tilt is None else (float(tilt[0]), float(tilt[1]))
    return build_scene(shot_config(depth, a, b, albedo, rough, sample_rate=sr, slope=sl, azimuth=az), hg_g=g)


def _scene_ids(spec, n):
    """Philox scene streams of a batch: scene_base + 0 .. n - 1 (spec
    "scene_base", default 0), so that groups of one scene set keep distinct
    streams."""
    b = int(spec.get("scene_base", 0))
    return range(b, b + n)


def _models(th, spec):
    tilts = spec.get("tilts")
    return [model_from_theta(row, spec.get("sample_rate"), None if tilts is None else tilts[i])
            for i, row in enumerate(th)]


def _replay_index(torch, sink, count, offsets, dev):
    L = runtime.load_library()
    keys = torch.empty(max(count, 1), dtype=torch.int64, device=dev)
    slots = torch.empty(max(count, 1), dtype=torch.int32, device=dev)
    if count == 0:
        return keys, slots, None
    b = C.c_size_t(0)
    runtime._check(L.alb_replay_index_workspace_size(count, C.byref(b)), "alb_replay_index_workspace_size")
    ws = torch.empty(b.value, dtype=torch.uint8, device=dev)
    runtime._check(L.alb_replay_index_build(runtime._p(sink.tensors["tag"]), runtime._p(sink.tensors["scene"]), count,
                                            runtime._p(offsets), runtime._p(keys), runtime._p(slots), runtime._p(ws),
                                            b.value, C.c_void_p(runtime._stream_ptr())), "alb_replay_index_build")
    return keys, slots, ws


def adjoint_weights(res, dloss):
    """Adjoint weights of a batched forward result kept with adjoint=True.

    dloss: flat float64 device tensor laid out like res.waveforms.  Returns
    (weights, direct): weights = (per-type record weights, per-type query
    weights), device tensors indexed by sink slot; direct [S, 3] = sums of
    w d log mult for (g, alpha^2, albedo * albedo)."""
    import torch

    st = res.state
    if not st:
        raise ValueError("simulate_batched(..., adjoint=True) keeps the state the adjoint needs")
    L = runtime.load_library()
    tab = st["table"]
    n = tab.n
    dev = res.waveforms.device
    s = C.c_void_p(runtime._stream_ptr())
    direct = torch.zeros(3 * n, dtype=torch.float64, device=dev)
    a_rec = [torch.zeros(max(st["rcounts"][t], 1), dtype=torch.float64, device=dev) for t in range(4)]
    a_q = [torch.zeros(max(st["qcounts"][t], 1), dtype=torch.float64, device=dev) for t in range(4)]
    dl = dloss.contiguous()
    for t, rw in st["rows"].items():
        m = st["maps"][t]
        q = st["queries"][t]
        runtime._check(L.alb_adjoint_contrib(runtime._p(tab.dev), n, tab.k, C.byref(m.struct), t, C.byref(q.struct),
                                             runtime._p(rw["order"]), 0, st["qcounts"][t], runtime._p(rw["nbr"]),
                                             runtime._p(rw["knorm"]), runtime._p(dl), runtime._p(st["wf_off"]),
                                             runtime._p(a_rec[t]), runtime._p(a_q[t]), runtime._p(direct), s),
                       "alb_adjoint_contrib")
    return (a_rec, a_q), direct.view(n, 3).cpu().numpy()


def replay_sums(res, weights, stat_layer: int = -1) -> np.ndarray:
    """Per-scene sums [S, 6] of w (n, [scatter], L, m, G, d) over records and
    queries, by path replay (alb_replay_batched): n omega0 factors, [scatter]
    a scatter record's own free-path density, L water length, m floor
    reflections, G phase-sampling score, d completed free paths ended by a
    scatter.  stat_layer = l restricts n, [scatter], L and d to layer l of
    layered columns (a single constant layer is layer 0); -1 counts all."""
    import torch

    st = res.state
    L = runtime.load_library()
    tab = st["table"]
    n = tab.n
    dev = res.waveforms.device
    s = C.c_void_p(runtime._stream_ptr())
    a_rec, a_q = weights
    acc = torch.zeros(REPLAY_SUMS * n, dtype=torch.float64, device=dev)
    ctr = torch.zeros(abi.N_COUNTERS, dtype=torch.int64, device=dev)
    keep = []
    for pass_, sinks, counts, wts, nph in ((abi.PASS_FORWARD, st["records"], st["rcounts"], a_rec, st["nf"]),
                                           (abi.PASS_BACKWARD, st["queries"], st["qcounts"], a_q, st["nb"])):
        offsets = torch.as_tensor(np.concatenate([[0], np.cumsum(nph)]).astype(np.int64), device=dev)
        idx = (abi.ReplayIndex * 4)()
        for t in range(4):
            key = ("replay_index", pass_, t)
            if key not in st:  # the sorted index of a sink serves every replay of this result
                st[key] = _replay_index(torch, sinks[t], counts[t], offsets, dev)
            keys, slots, _ = st[key]
            idx[t].keys = runtime._p(keys)
            idx[t].slots = runtime._p(slots)
            idx[t].n = counts[t] if t in st["rows"] else 0
            idx[t].weight = runtime._p(wts[t])
        rng = abi.make_rng(st["key"], pass_ | (int(st["base"]) << 1))
        work = torch.zeros(1, dtype=torch.int64, device=dev)
        runtime._check(L.alb_replay_batched(runtime._p(tab.dev), n, runtime._p(offsets), C.byref(rng), pass_,
                                            C.cast(idx, C.c_void_p), runtime._p(work), st["block"], st["refill"],
                                            runtime._p(acc), runtime._p(ctr), int(stat_layer), s),
                       "alb_replay_batched")
        keep.append((offsets, work))
    misses = int(ctr[abi.CTR_REPLAY_MISSES].item())
    if misses:
        raise runtime.AlbError(f"path replay did not find {misses} records")
    del keep
    return acc.view(n, REPLAY_SUMS).cpu().numpy()


def adjoint_sums(res, dloss):
    """Per-scene sums from a batched forward result kept with adjoint=True.

    dloss: flat float64 device tensor laid out like res.waveforms.  Returns
    (acc [S, 6] = replay_sums over the whole column, direct [S, 3] = sums of
    w d log mult for (g, alpha^2, albedo * albedo)).
    """
    weights, direct = adjoint_weights(res, dloss)
    return replay_sums(res, weights), direct


def gradients_from_sums(acc, direct, derivs) -> np.ndarray:
    """theta-gradients [S, 6] (depth column left at 0) from the per-scene sums."""
    out = np.zeros((len(derivs), 6))
    for s, d in enumerate(derivs):
        c0, cs, c2, c3, c4, c5 = acc[s]
        alpha, a = d.alpha, d.a
        # free-path densities (alpha e^{-alpha L}) and omega0 = (alpha - a) / alpha factors
        dl_dalpha = (cs + c5) / alpha + c0 * (1.0 / (alpha - a) - 1.0 / alpha) - c2
        out[s, 0] = dl_dalpha * d.dalpha[0] - c0 / (alpha - a)
        out[s, 1] = dl_dalpha * d.dalpha[1]
        out[s, 2] = dl_dalpha * d.dalpha[2] + c4 + direct[s, 0]
        out[s, 4] = (c3 + direct[s, 2]) / d.albedo
        out[s, 5] = direct[s, 1] * 4.0 * d.roughness ** 3  # alpha^2 = r^4
    return out


def layer_derivatives(model):
    """Per layer of a layered column of constant layers: (alpha, a,
    d alpha / d a, d alpha / d b), alpha the float32 lidar attenuation the
    kernels use, by central differences of the Churnside formula with the
    layer's salinity and backscatter fraction."""
    cfg = model.config
    H = cfg.scene.flying_height
    lam = cfg.laser.wavelength
    div = cfg.laser.divergence_angle
    out = []
    for layer, lm in zip(cfg.water.layers, model.layers):
        a, b = float(layer.absorption_coefficient), float(layer.scattering_coefficient)

        def al(aa, bb):
            return lidar_attenuation(aa, bb, layer.salinity, lam, H, div, lm.backscatter)

        ha, hb = 1e-6 * max(abs(a), 1e-3), 1e-6 * max(abs(b), 1e-3)
        out.append((float(np.float32(lm.alpha64)), a, (al(a + ha, b) - al(a - ha, b)) / (2 * ha),
                    (al(a, b + hb) - al(a, b - hb)) / (2 * hb)))
    return out


def layered_gradients(res, dloss, models):
    """Loss gradients of scenes whose water column is a stack of constant
    layers (the reference's TurbidityLayerConfig list, physics/models/
    water.py:253-416), from a batched forward result kept with adjoint=True.

    Returns one dict per scene: "a" and "b" (per layer, d loss / d
    absorption and scattering coefficient), "albedo" and "roughness".  One
    replay per layer restricts the water length and scatter statistics to
    that layer; omega0 counts at the scatter's new position and free-path
    densities at the sub-step's layer, as the trace applies them.  A scene
    with a single constant layer gives layer 0 (same values as
    gradients_from_sums).  The gather's phase mass is the layer's fixed
    Fournier-Forand table, so there is no phase parameter."""
    for m in models:
        if m.layers and any(lm.profiled for lm in m.layers):
            raise NotImplementedError("layered gradients need constant layers (no depth profiles)")
    weights, direct = adjoint_weights(res, dloss)
    n_layers = max(len(m.layers) if m.layers else 1 for m in models)
    sums = [replay_sums(res, weights, l) for l in range(n_layers)]
    out = []
    for s, m in enumerate(models):
        derivs = layer_derivatives(m) if m.layers else [(float(m.alpha_l), float(m.absorption),
                                                          *_single_layer_dalpha(m))]
        out.append(layer_gradients_from_sums([sums[l][s] for l in range(len(derivs))], direct[s], derivs,
                                             m.floor_albedo, m.roughness))
    return out


def layer_gradients_from_sums(layer_sums, direct, derivs, albedo: float, roughness: float) -> dict:
    """One scene's gradients from its per-layer replay sums (layer_sums[l] =
    replay_sums(..., stat_layer=l)[scene]), its direct sums and, per layer,
    (alpha, a, d alpha / d a, d alpha / d b):
    d L / d alpha_l = (d_l + [scatter]_l) / alpha_l + n_l (1 / (alpha_l - a_l)
    - 1 / alpha_l) - L_l, d L / d a_l = d L / d alpha_l d alpha_l / d a_l
    - n_l / (alpha_l - a_l), d L / d b_l = d L / d alpha_l d alpha_l / d b_l;
    the floor reflections m and the direct sums are whole-path terms."""
    ga, gb = [], []
    for (c0, cs, c2, _, _, c5), (alpha, a, da_a, da_b) in zip(layer_sums, derivs):
        dl_dalpha = (cs + c5) / alpha + c0 * (1.0 / (alpha - a) - 1.0 / alpha) - c2
        ga.append(dl_dalpha * da_a - c0 / (alpha - a))
        gb.append(dl_dalpha * da_b)
    return {"a": np.array(ga), "b": np.array(gb), "albedo": (layer_sums[0][3] + direct[2]) / albedo,
            "roughness": direct[1] * 4.0 * roughness ** 3}


def _single_layer_dalpha(m):
    layer = m.config.water.layers[0]
    cfg = m.config
    a, b = float(layer.absorption_coefficient), float(layer.scattering_coefficient)
    bp = m.backscatter

    def al(aa, bb):
        return lidar_attenuation(aa, bb, layer.salinity, cfg.laser.wavelength, cfg.scene.flying_height,
                                 cfg.laser.divergence_angle, bp)

    ha, hb = 1e-6 * max(abs(a), 1e-3), 1e-6 * max(abs(b), 1e-3)
    return (al(a + ha, b) - al(a - ha, b)) / (2 * ha), (al(a, b + hb) - al(a, b - hb)) / (2 * hb)


class BatchedWaveformFn:
    """torch.autograd.Function over the batched form: theta [S, 6] (float64)
    -> gated waveforms [S, 4, win_len] (float64, on the GPU)."""

    @staticmethod
    def build():
        import torch

        class _Fn(torch.autograd.Function):
            @staticmethod
            def forward(ctx, theta, spec):
                from .execution.batched import simulate_batched

                th = theta.detach().cpu().numpy().astype(np.float64)
                models = _models(th, spec)
                gr = spec.get("graphs")
                res = simulate_batched(models, spec["n_forward"], spec["n_backward"], spec["key"],
                                       scene_ids=_scene_ids(spec, len(models)),
                                       win_len=spec.get("win_len", 1000), strict=False, adjoint=True,
                                       block=spec.get("block", 256), refill=spec.get("refill", 4),
                                       graphs=None if gr is None else gr.setdefault("forward", {}))
                wl = spec.get("win_len", 1000)
                out = torch.stack([res.gated(s, models[s].gate(wl)) for s in range(len(models))])
                ctx.res = res
                ctx.out = out.detach()  # the base of the one-sided depth difference
                ctx.models = models
                ctx.theta = th
                ctx.spec = spec
                ctx.device = theta.device
                return out

            @staticmethod
            def backward(ctx, grad_out):
                res, models, th, spec = ctx.res, ctx.models, ctx.theta, ctx.spec
                wl = spec.get("win_len", 1000)
                dloss = torch.zeros_like(res.waveforms)
                go = grad_out.to(res.waveforms.device, torch.float64)
                for s, m in enumerate(models):
                    g0, gl = m.gate(wl)
                    o = int(res.offsets[s])
                    nb = m.n_bins
                    dloss[o:o + 4 * nb].view(4, nb)[:, g0:g0 + gl] = go[s]
                acc, direct = adjoint_sums(res, dloss)
                # the kept forward state is no longer needed: free it before
                # the depth difference runs its own forward pass
                ctx.res = None
                del res, dloss
                derivs = [scene_derivatives(m, th[s]) for s, m in enumerate(models)]
                grads = gradients_from_sums(acc, direct, derivs)
                mode = spec.get("depth_grad", "fd1")
                if mode == "fd":
                    grads[:, 3] = depth_fd(th, spec, go, spec.get("depth_rel", 1e-3))
                elif mode == "fd1":
                    grads[:, 3] = depth_fd(th, spec, go, spec.get("depth_rel", 1e-3), base=ctx.out)
                ctx.out = None
                ctx.res = None
                return torch.as_tensor(grads, dtype=torch.float64, device=ctx.device), None

        return _Fn


def depth_fd(th, spec, go, rel: float = 1e-3, base=None):
    """d loss / d depth by differences with common random numbers: central
    (two more forward passes), or one-sided from base, the gated waveforms
    already computed at th with the same key (one more forward pass)."""
    import torch

    from .execution.batched import simulate_batched

    wl = spec.get("win_len", 1000)
    outs = []
    for sign in ((1.0,) if base is not None else (1.0, -1.0)):
        t2 = th.copy()
        t2[:, 3] = th[:, 3] * (1.0 + sign * rel)
        models = _models(t2, spec)
        gr = spec.get("graphs")
        r = simulate_batched(models, spec["n_forward"], spec["n_backward"], spec["key"], win_len=wl, strict=False,
                             scene_ids=_scene_ids(spec, len(models)),
                             block=spec.get("block", 256), refill=spec.get("refill", 4),
                             graphs=None if gr is None else gr.setdefault("depth_difference", {}))
        outs.append(torch.stack([r.gated(s, models[s].gate(wl)) for s in range(len(models))]))
    if base is not None:
        d = (outs[0] - base.to(outs[0].device)) * go
        return (d.sum(dim=(1, 2)) / torch.as_tensor(rel * th[:, 3], device=d.device)).cpu().numpy()
    d = (outs[0] - outs[1]) * go
    return (d.sum(dim=(1, 2)) / torch.as_tensor(2.0 * rel * th[:, 3], device=d.device)).cpu().numpy()


def gated_waveforms(theta, n_forward, n_backward, key, **spec):
    """Gated waveforms [S, 4, win_len] of the scenes theta [S, 6] without
    what the adjoint needs (no kept sinks or neighbour rows): for values only."""
    import torch

    from .execution.batched import simulate_batched

    th = theta.detach().cpu().numpy().astype(np.float64) if hasattr(theta, "detach") else np.asarray(theta)
    models = _models(th, spec)
    wl = spec.get("win_len", 1000)
    res = simulate_batched(models, n_forward, n_backward, key, win_len=wl, strict=False,
                           scene_ids=_scene_ids(spec, len(models)),
                           block=spec.get("block", 256), refill=spec.get("refill", 4))
    return torch.stack([res.gated(s, models[s].gate(wl)) for s in range(len(models))])


_FN = None


def batched_waveforms(theta, n_forward, n_backward, key, **spec):
    """Differentiable gated waveforms [S, 4, win_len] of the scenes theta [S, 6]."""
    global _FN
    if _FN is None:
        _FN = BatchedWaveformFn.build()
    spec = dict(spec, n_forward=n_forward, n_backward=n_backward, key=key)
    return _FN.apply(theta, spec)


def layered_config(base, row):
    """base (a SimulationConfig whose water column is a stack of constant
    layers) with the layers' absorption and scattering coefficients and the
    floor albedo and surface roughness taken from row = (a_0 .. a_{L-1},
    b_0 .. b_{L-1}, albedo, roughness)."""
    import dataclasses

    from .config import SeaFloorConfig, SeaSurfaceConfig, WaterConfig

    layers = base.water.layers
    nl = len(layers)
    row = [float(v) for v in row]
    if len(row) != 2 * nl + 2:
        raise ValueError(f"expected {2 * nl + 2} parameters for {nl} layers, got {len(row)}")
    new = tuple(dataclasses.replace(ly, absorption_coefficient=row[i], scattering_coefficient=row[nl + i])
                for i, ly in enumerate(layers))
    return dataclasses.replace(base, water=dataclasses.replace(base.water, layers=new),
                               sea_floor=dataclasses.replace(base.sea_floor, albedo=row[2 * nl]),
                               sea_surface=dataclasses.replace(base.sea_surface, roughness=row[2 * nl + 1]))


class LayeredWaveformFn:
    """torch.autograd.Function over the batched form for layered columns:
    params [S, 2 L + 2] (float64; per scene (a_0 .. a_{L-1}, b_0 .. b_{L-1},
    albedo, roughness) of the base configuration spec["configs"][s], all with
    L constant layers) -> gated waveforms [S, 4, win_len]; the backward pass
    is layered_gradients (path replay, one replay per layer)."""

    @staticmethod
    def build():
        import torch

        from .model import build_scene

        class _Fn(torch.autograd.Function):
            @staticmethod
            def forward(ctx, params, spec):
                from .execution.batched import simulate_batched

                p = params.detach().cpu().numpy().astype(np.float64)
                models = [build_scene(layered_config(cfg, p[s])) for s, cfg in enumerate(spec["configs"])]
                wl = spec.get("win_len", 1000)
                res = simulate_batched(models, spec["n_forward"], spec["n_backward"], spec["key"],
                                       scene_ids=_scene_ids(spec, len(models)), win_len=wl, strict=False,
                                       adjoint=True, block=spec.get("block", 256), refill=spec.get("refill", 4))
                ctx.res, ctx.models, ctx.wl, ctx.device = res, models, wl, params.device
                return torch.stack([res.gated(s, models[s].gate(wl)) for s in range(len(models))])

            @staticmethod
            def backward(ctx, grad_out):
                res, models, wl = ctx.res, ctx.models, ctx.wl
                dloss = torch.zeros_like(res.waveforms)
                go = grad_out.to(res.waveforms.device, torch.float64)
                for s, m in enumerate(models):
                    g0, gl = m.gate(wl)
                    o, nb = int(res.offsets[s]), m.n_bins
                    dloss[o:o + 4 * nb].view(4, nb)[:, g0:g0 + gl] = go[s]
                g = layered_gradients(res, dloss, models)
                ctx.res = None
                rows = [np.concatenate([x["a"], x["b"], [x["albedo"], x["roughness"]]]) for x in g]
                return torch.as_tensor(np.array(rows), dtype=torch.float64, device=ctx.device), None

        return _Fn


_LFN = None


def layered_waveforms(params, configs, n_forward, n_backward, key, **spec):
    """Differentiable gated waveforms [S, 4, win_len] of layered columns:
    params [S, 2 L + 2] = (a_0 .. a_{L-1}, b_0 .. b_{L-1}, albedo, roughness)
    applied to the base configurations configs[s]."""
    global _LFN
    if _LFN is None:
        _LFN = LayeredWaveformFn.build()
    spec = dict(spec, configs=list(configs), n_forward=n_forward, n_backward=n_backward, key=key)
    return _LFN.apply(params, spec)
