"""The uneven-size ensemble at its real sizes: the 64 FP32 free-convection members of
ensemble_members(seed=5) (sizes from {64³, 96×96×64, 128×128×64, 128³}, each with its own noise,
N² and surface buoyancy flux) advanced together in the batched form (one launch per kernel of the
step over a table of members, members in four groups whose graphs run concurrently, one replay
per pair of steps) for 20 steps, through the C ABI.

Checks: every member's state stays finite and projected (max |∇·U| at FP32 rounding of the
velocity update, relative to max |U| / Δx); one member of each size, twinned and stepped on its
own with om_time_step, equals its batched self bit for bit."""
import numpy as np
import pytest

import oceananigans_mi355x as om
from cases import product_for
from oceananigans_mi355x import workloads

pytestmark = pytest.mark.gpu


def test_sixty_four_uneven_members_batched_at_full_size():
    cfgs = workloads.ensemble_members(n=64, seed=5)
    members = [product_for(*c) for c in cfgs]
    first = {}
    for q, (grid, _, _) in enumerate(cfgs):
        first.setdefault(grid.N, q)
    twins = {q: product_for(*cfgs[q]) for q in first.values()}
    print("sizes:", {"x".join(map(str, k)): sum(1 for g, _, _ in cfgs if g.N == k) for k in first}, flush=True)
    ens = om.Ensemble(members)
    ens.set_batch_groups(4)
    ens.time_step(1.0, 20, mode="batched")
    for q, t in twins.items():
        t.time_step(1.0, 20)
        m = members[q]
        for n in ("u", "v", "w", "b"):
            a = (m.velocities[n] if n in "uvw" else m.tracers[n]).interior()
            b = (t.velocities[n] if n in "uvw" else t.tracers[n]).interior()
            np.testing.assert_array_equal(a, b, err_msg=f"member {q} {n}")
        assert m.clock._get() == t.clock._get()
    worst = 0.0
    for q, m in enumerate(members):
        for f in (m.velocities.u, m.velocities.v, m.velocities.w, m.tracers.b):
            assert not f.hasnan(), (q, f.id)
        div = np.abs(m.divergence()).max()
        umax = max(np.abs(m.velocities[n].interior()).max() for n in "uvw")
        dx = m.grid.extent[0] / m.grid.N[0]
        worst = max(worst, div * dx / umax)
        assert m.clock.iteration == 20
    print("max over members of max|div U| dx / max|U|:", worst, flush=True)
    assert worst < 1e-4
    ens.close()
