"""The batched annual sweep at its full size: the 100-paraboloid field under
512 suns (32 azimuths x 16 elevations, 5-85 degrees) with seeded
log-uniform counts in [1e5, 1e7] (1.24e9 realisations), one
ssol_solve_sun_batch call.

- Every sun: its count honoured, and its sinks plus shadowing balance its
  initial flux. Each realisation's sinks sum to its w0 exactly
  (src/ssol_solver.c:551-557), and E[w0] = DNI x sampled area x cos factor
  up to the sun's angular blur (sigma 2.51 mrad: below 1e-4 relative).
- Four suns across the sweep against single-sun solves of the oracle (the
  same field under that sun): every global estimate within 3 combined
  standard errors (src/test_ssol_solver6.c:219-235 merge semantics: batch
  sun j is a solve of counts[j] realisations)."""
import dataclasses

import numpy as np
import pytest

import oracle_ctypes as O
import ssol_ctypes as L
import ssol_scene as S
from conftest import within

pytestmark = pytest.mark.gpu

SINKS = ("absorbed_by_receivers", "missing", "shadowed", "extinguished_by_atmosphere",
         "other_absorbed")


@pytest.fixture(scope="module")
def sweep(gpu):
    spec, dirs, counts = S.config4()
    assert len(dirs) == 512 and 1.1e9 < counts.sum() < 1.4e9
    sc = L.Scene(spec)
    area = sc.solve(10_000, seed=1).sampled_area
    res, got = sc.solve_sun_batch(dirs, counts, seed=21)
    sc.close()
    return spec, dirs, counts, res, got, area


def test_sweep_counts_and_energy_balance_every_sun(sweep):
    spec, dirs, counts, res, got, area = sweep
    assert list(got) == [int(c) for c in counts]
    worst = 0.0
    for j, g in enumerate(res):
        sinks = sum(g[f][0] for f in SINKS)
        initial = spec.sun.dni * area * g["cos_factor"][0]
        assert initial > 0, j
        rel = abs(sinks - initial) / initial
        worst = max(worst, rel)
        assert rel <= 1e-4, (j, sinks, initial)
    print(f"largest energy imbalance over 512 suns: {worst:.3g} relative")


@pytest.mark.parametrize("j", [3, 170, 341, 508])
def test_sweep_suns_vs_oracle(sweep, j):
    spec, dirs, counts, res, got, area = sweep
    one = dataclasses.replace(spec, sun=dataclasses.replace(spec.sun, direction=tuple(dirs[j])),
                              name=f"config4_sun{j}")
    o = O.Oracle(one, track_sampled_x_receiver=False)
    n = 1_000_000
    assert o.solve(n, seed=77) == 0
    og = o.global_()
    bad = []
    for f in L.GLOBAL_FIELDS:
        ok, tol = within(res[j][f], og[f], 3.0, floor=1e-9 * max(1.0, abs(og[f][0])))
        if not ok:
            bad.append((f, res[j][f], og[f], tol))
    print(f"sun {j} ({int(counts[j])} paths): absorbed {res[j]['absorbed_by_receivers'][0]:.6g} "
          f"+- {res[j]['absorbed_by_receivers'][2]:.3g}, oracle {og['absorbed_by_receivers'][0]:.6g} "
          f"+- {og['absorbed_by_receivers'][2]:.3g}")
    assert not bad, bad
