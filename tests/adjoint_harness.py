"""Differentiable-aiming harness: the GPU forward solve and its hand-written
adjoint (ssol_solve_adjoint) wrapped as a torch.autograd.Function, and
common-random-number finite differences of the GPU solver and of the oracle.

Parameters of heliostat i: aim angles a[i, 0], a[i, 1] (the heliostat's
rotation is R rot_x(a0) rot_y(a1), R the nominal aiming) and slope error
s[i] (Beckmann roughness sqrt(2) s[i]). Outputs: absorbed_by_receivers (W)
and the receiver's front-side per-triangle absorbed flux map (W/m2)."""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

import ssol_ctypes as L
import ssol_scene as S


@dataclass
class Field:
    seed: int = 0
    n_helio: int = 30
    radius: float = 300.0
    rcv_size: float = 3.0
    rcv_cells: int = 16
    n: int = 4_000_000
    rng_seed: int = 1
    kw: dict = field(default_factory=dict)
    path_slots: int = 0        # wavefront width of the GPU solves (0: default)
    sample_only: int = -1      # >= 0: only this heliostat is sampled

    def scene(self, aim, slope):
        spec = S.config2(self.seed, n_helio=self.n_helio, radius=self.radius,
                         aim_perturb=np.asarray(aim), slope_errors=list(np.asarray(slope)),
                         rcv_size=self.rcv_size, rcv_cells=self.rcv_cells, **self.kw)
        if self.sample_only >= 0:
            for j, inst in enumerate(spec.instances[:-1]):
                inst.sample = 1 if j == self.sample_only else 0
        axes = S.aim_axes(self.seed, self.n_helio, self.radius, self.kw.get("sun_dir"),
                          np.asarray(aim))
        return spec, axes


def receiver_of(spec):
    rcv = spec.instances[-1]
    return rcv, rcv.object.sshapes[0][0]


def forward(fld: Field, aim, slope, seed=None, n=None):
    """(absorbed, map) of the GPU solve."""
    spec, _ = fld.scene(aim, slope)
    lib = L.Scene(spec, path_slots=fld.path_slots)
    est = lib.solve(n or fld.n, seed=fld.rng_seed if seed is None else seed)
    absorbed = est.global_()["absorbed_by_receivers"][0]
    rcv, shape = receiver_of(spec)
    fmap = est.primitives(rcv, shape, S.FRONT, "absorbed_flux")[:, 0].copy()
    lib.close()
    return absorbed, fmap


def adjoint(fld: Field, aim, slope, d_absorbed, d_map, seed=None, n=None):
    """Gradient [n_helio][3] (d/d a0, d/d a1, d/d slope error)."""
    spec, axes = fld.scene(aim, slope)
    lib = L.Scene(spec, path_slots=fld.path_slots)
    rcv, shape = receiver_of(spec)
    g = lib.solve_adjoint(n or fld.n, spec.instances[:-1], axes, d_absorbed=d_absorbed,
                          map_receiver=rcv if d_map is not None else None,
                          map_shape=shape if d_map is not None else None, d_map=d_map,
                          seed=fld.rng_seed if seed is None else seed)
    lib.close()
    return g


class FluxFunction(torch.autograd.Function):
    """(aim [n][2], slope [n]) -> (absorbed, flux map) with the adjoint as
    backward; the same realisations (seed) serve both passes."""

    @staticmethod
    def forward(ctx, aim, slope, fld):
        ctx.fld = fld
        ctx.save_for_backward(aim, slope)
        a, m = forward(fld, aim.detach().cpu().numpy(), slope.detach().cpu().numpy())
        return (torch.tensor(a, dtype=torch.float64),
                torch.tensor(m, dtype=torch.float64))

    @staticmethod
    def backward(ctx, g_abs, g_map):
        aim, slope = ctx.saved_tensors
        d_map = None if g_map is None else g_map.detach().cpu().numpy().astype(np.float64)
        d_abs = 0.0 if g_abs is None else float(g_abs)
        g = adjoint(ctx.fld, aim.detach().cpu().numpy(), slope.detach().cpu().numpy(), d_abs,
                    d_map)
        return (torch.tensor(g[:, :2], dtype=aim.dtype), torch.tensor(g[:, 2], dtype=slope.dtype),
                None)


def map_variance(fmap):
    return ((fmap - fmap.mean()) ** 2).mean()


def crn_fd(fld: Field, aim, slope, i, k, h, objective, seed=None, n=None, solver="gpu"):
    """Central difference of objective(absorbed, map) in parameter (i, k)
    (k = 0, 1: aim angles; 2: slope error), same random numbers both sides."""
    vals = []
    for sgn in (+1, -1):
        a = np.array(aim, dtype=np.float64, copy=True)
        s = np.array(slope, dtype=np.float64, copy=True)
        if k < 2:
            a[i, k] += sgn * h
        else:
            s[i] += sgn * h
        if solver == "gpu":
            ab, fm = forward(fld, a, s, seed=seed, n=n)
        else:
            ab, fm = forward_oracle(fld, a, s, seed=seed, n=n)
        vals.append(objective(ab, fm))
    return (vals[0] - vals[1]) / (2 * h)


def forward_oracle(fld: Field, aim, slope, seed=None, n=None):
    import oracle_ctypes as O
    spec, _ = fld.scene(aim, slope)
    o = O.Oracle(spec, track_sampled_x_receiver=False)
    assert o.solve(n or fld.n, seed=fld.rng_seed if seed is None else seed) == 0
    absorbed = o.global_()["absorbed_by_receivers"][0]
    rcv, shape = receiver_of(spec)
    s_idx = 0
    nt = o.shape_ntris(shape)
    fmap = np.array([o.primitive(rcv, s_idx, t, S.FRONT)["absorbed_flux"][0] for t in range(nt)])
    # the oracle's per-primitive values are per triangle area like the getter
    return absorbed, fmap
