"""Checks of recorded paths (the reference's path tracker,
src/ssol_solver.c:898-1127, read back through src/ssol_estimator.c:270-322),
shared by the oracle tests and the GPU tests.

A path is (type, V) with V[k] = (x, y, z, weight). With a sun length > 0 the
first vertex lies sun_length upstream of the start point and carries the
DNI; the start point carries the initial flux w0; every traced bounce ray
adds one vertex: its hit point, or, when it escapes and the escape length is
> 0, the point escape_length beyond the last point, weighted by the flux
that left that point. These rules give exact identities with the tallies of
the same solve:
  sum over shadow paths of w0 / N           = shadowed
  sum over escaped paths of last weight / N = missing
  sum over paths of (vertices - 2)          = traced rays - N
(the last one when no path was dropped by the roulette)."""
from __future__ import annotations

import math

import numpy as np

import ssol_scene as S


def two_mirrors(gap: float = 0.11, theta_deg: float = 45.0) -> S.Scene:
    """Two parallel specular mirrors (reflectivity 1) facing each other,
    `gap` apart, lit through an opening: the lower mirror (z = 0) is a
    sampled entrance strip over x in [1.9, 2.1] and an unsampled part over
    x in [2.1, 10]; the upper one (z = gap, facing down) spans x in [2, 12].
    Under a directional sun at theta_deg from the vertical every start point
    is lit and its reflection enters the gap, then advances gap tan(theta)
    per segment: about 8 / (gap tan theta) segments per path, so each path
    meets the roulette at depth 65."""
    t = math.radians(theta_deg)
    sun_dir = (math.sin(t), 0.0, -math.cos(t))
    m = S.mirror(1.0)
    entrance = S.Instance(S.Object([(S.rect_mesh(1.9, 2.1, -1.0, 1.0), m, m)]), S.IDENTITY,
                          receiver_mask=0, sample=1)
    lower = S.Instance(S.Object([(S.rect_mesh(2.1, 10.0, -1.0, 1.0), m, m)]), S.IDENTITY,
                       receiver_mask=0, sample=0)
    upper = S.Instance(S.Object([(S.rect_mesh(2.0, 12.0, -1.0, 1.0), m, m)]),
                       S.transform(S.rot_x(math.pi), (0.0, 0.0, gap)), receiver_mask=0,
                       sample=0)
    sun = S.Sun("directional", sun_dir, 1000.0, S.Spectrum([1.0, 2.0, 3.0], [1.0, 0.8, 1.0]))
    return S.Scene([entrance, lower, upper], sun, None, name=f"two_mirrors_{gap}")


def solver1_extent() -> float:
    """0.75 x the largest extent of src/test_ssol_solver1.c's scene box: unit
    half-size squares at the origin, flipped about X at (2, 0, 2) and at
    (4, 0, 0): x in [-1, 5], y in [-1, 1], z in [0, 2]."""
    return 0.75 * 6.0


def check_paths(paths, n, dni, sun_len, inf_len, g, segments=None, rtol=1e-9):
    """Structure of every path and the identities above; returns the number
    of escaped paths."""
    start = 1 if sun_len > 0 else 0
    shadow_sum = missing_sum = 0.0
    extra = 0
    escaped = 0
    for ty, v in paths:
        assert ty in ("missing", "shadow", "success", "error", 0, 1, 2, 3), ty
        ty = ("missing", "shadow", "success", "error")[ty] if isinstance(ty, int) else ty
        assert len(v) >= start + 1, (ty, v)
        if sun_len > 0:
            assert v[0, 3] == dni
            d = np.linalg.norm(v[1, :3] - v[0, :3])
            assert abs(d - sun_len) <= 1e-9 * sun_len, (d, sun_len)
        assert v[start, 3] > 0  # initial flux
        if ty == "shadow":
            assert len(v) == start + 1, v
            shadow_sum += v[start, 3]
        elif inf_len > 0 and len(v) >= start + 2:
            d = np.linalg.norm(v[-1, :3] - v[-2, :3])
            if abs(d - inf_len) <= 1e-9 * inf_len:
                escaped += 1
                missing_sum += v[-1, 3]
        extra += len(v) - start - 1
    scale = max(abs(g["shadowed"][0]), 1.0)
    assert abs(shadow_sum / n - g["shadowed"][0]) <= rtol * scale, (shadow_sum / n, g["shadowed"])
    if inf_len > 0:
        scale = max(abs(g["missing"][0]), 1.0)
        assert abs(missing_sum / n - g["missing"][0]) <= rtol * scale, (missing_sum / n, g["missing"])
    if segments is not None and len(paths) == n and inf_len > 0:
        assert extra == segments - n, (extra, segments, n)
    return escaped


def check_alternating(paths, gap, sun_len, inf_len):
    """Lit paths of two_mirrors: after the start point on the lower mirror
    every hit alternates between z = gap and z = 0, and the last vertex is
    the escape point."""
    start = 1 if sun_len > 0 else 0
    for ty, v in paths:
        if ty in (1, "shadow"):
            continue
        hits = v[start:-1] if inf_len > 0 else v[start:]
        assert abs(hits[0, 2]) < 1e-9
        for k in range(1, len(hits)):
            want = gap if k % 2 == 1 else 0.0
            assert abs(hits[k, 2] - want) < 1e-6, (k, hits[k], want)


def type_counts(paths):
    names = ("missing", "shadow", "success", "error")
    out = dict.fromkeys(names, 0)
    for ty, _ in paths:
        out[names[ty] if isinstance(ty, int) else ty] += 1
    return out
