"""Ensembles of independent models advanced together and the z-chunk autotuner, through the C ABI:
member graphs (each member's two-step launch sequence replayed on its own stream) and the batched
form (one launch per kernel of the step over a table of all members, the whole sequence one
hipGraph).

Both forms run the same kernel bodies with the same arguments as the direct launches, so every
member must match a twin advanced by om_time_step bit for bit, after each of several replays with
the program's own eager work and synchronisations in between; the batched members also match the
FP64 oracle."""
import numpy as np
import pytest

import oceananigans_mi355x as om
from cases import convection, product_for, seamount

pytestmark = pytest.mark.gpu

NAMES = ("u", "v", "w", "b")


def _members():
    cfgs = [convection(N=(16, 16, 16), float_type=np.float32, noise_velocity=1e-2, seed=1),
            convection(N=(24, 20, 12), float_type=np.float32, noise_velocity=1e-2, seed=2),
            convection(N=(16, 16, 8), float_type=np.float64, noise_velocity=1e-2, seed=3)]
    return [product_for(*c) for c in cfgs], [product_for(*c) for c in cfgs]


def _same(a, b):
    for n in NAMES:
        fa = a.velocities[n] if n in "uvw" else a.tracers[n]
        fb = b.velocities[n] if n in "uvw" else b.tracers[n]
        np.testing.assert_array_equal(fa.interior(), fb.interior(), err_msg=n)
    assert a.clock._get() == b.clock._get()


@pytest.mark.parametrize("graphs", [True, False])
def test_ensemble_steps_match_members_stepped_one_after_another(graphs):
    members, twins = _members()
    ens = om.Ensemble(members)
    for nsteps in (2, 3, 2, 1, 4):          # odd counts flip the G⁻/Gⁿ parity between replays
        ens.time_step(0.5, nsteps, graphs=graphs)
        for t in twins:
            t.time_step(0.5, nsteps)
        for m, t in zip(members, twins):
            _same(m, t)
        print("steps", nsteps, "iterations", [m.clock.iteration for m in members], flush=True)
    ens.close()


def test_changed_dt_and_tuning_recapture():
    members, twins = _members()
    ens = om.Ensemble(members)
    ens.time_step(0.5, 4)
    members[0].set_tuning(kchunk=5)         # drops the member's captured steps
    ens.time_step(0.25, 4)                  # a new Δt is captured anew
    for t in twins:
        t.time_step(0.5, 4)
        t.time_step(0.25, 4)
    for m, t in zip(members, twins):
        _same(m, t)
    ens.close()


def test_autotune_picks_a_chunk_and_leaves_the_state():
    members, twins = _members()
    m, t = members[0], twins[0]
    m.time_step(0.5, 1)
    t.time_step(0.5, 1)
    g_before = m.Gn["u"].interior().copy()
    kc = m.autotune()
    print("autotuned kchunk", kc, flush=True)
    assert kc in (4, 8, 12, 16)
    assert m.get_tuning()["kchunk"] == kc
    np.testing.assert_array_equal(m.Gn["u"].interior(), g_before)
    m.time_step(0.5, 3)
    t.time_step(0.5, 3)
    _same(m, t)


def test_conjugate_gradient_members_are_rejected():
    grid, kw, init = seamount(N=(16, 16, 12))
    cg = product_for(grid, kw, init)
    with pytest.raises(om.OmError, match="conjugate-gradient"):
        om.Ensemble([cg])


def _same_precision_members(float_type, seeds=(11, 12, 13)):
    cfgs = [convection(N=(16, 16, 16), float_type=float_type, noise_velocity=1e-2, seed=seeds[0]),
            convection(N=(24, 20, 12), float_type=float_type, noise_velocity=1e-2, seed=seeds[1]),
            convection(N=(32, 16, 8), float_type=float_type, noise_velocity=1e-2, seed=seeds[2])]
    return cfgs, [product_for(*c) for c in cfgs], [product_for(*c) for c in cfgs]


@pytest.mark.parametrize("float_type", [np.float32, np.float64])
def test_batched_ensemble_matches_members_stepped_one_after_another(float_type):
    _, members, twins = _same_precision_members(float_type)
    ens = om.Ensemble(members)
    for nsteps in (1, 2, 3, 2, 4, 1):      # the first step runs member by member; odd counts leave a remainder
        ens.time_step(0.5, nsteps, mode="batched")
        for t in twins:
            t.time_step(0.5, nsteps)
        for m, t in zip(members, twins):
            _same(m, t)
        print(float_type.__name__, "batched steps", nsteps, "iterations", [m.clock.iteration for m in members], flush=True)
    members[1].set_tuning(kchunk=3)           # a member's launch shape changes: the tables are rebuilt
    ens.time_step(0.25, 4, mode="batched")
    twins[1].set_tuning(kchunk=3)
    for t in twins:
        t.time_step(0.25, 4)
    for m, t in zip(members, twins):
        _same(m, t)
    ens.set_batch_groups(2)                   # two member-table graphs on concurrent streams
    ens.time_step(0.25, 6, mode="batched")
    for t in twins:
        t.time_step(0.25, 6)
    for m, t in zip(members, twins):
        _same(m, t)
    with pytest.raises(om.OmError):
        ens.set_batch_groups(0)
    ens.close()


def test_batched_members_match_the_oracle():
    from cases import oracle_for, rel_err
    cfgs, members, _ = _same_precision_members(np.float64)
    ens = om.Ensemble(members)
    ens.time_step(1.0, 5, mode="batched")
    for (grid, kw, init), m in zip(cfgs, members):
        o = oracle_for(grid, kw, init)
        o.time_step(1.0, 5)
        for q, n in enumerate(NAMES):
            f = m.velocities[n] if q < 3 else m.tracers[n]
            e = rel_err(f.interior(), o.get(q))
            print("member", grid.N, n, "rel err vs oracle", e, flush=True)
            assert e <= 1e-11
    ens.close()


def test_batched_ensemble_rejects_members_with_other_launch_sequences():
    members, _ = _members()                   # FP32 and FP64 members
    ens = om.Ensemble(members)
    with pytest.raises(om.OmError, match="launch sequence"):
        ens.time_step(0.5, 4, mode="batched")
    ens.close()
