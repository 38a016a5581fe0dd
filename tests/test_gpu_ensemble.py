"""Ensembles of independent runs (coral_amd_ensemble_*): one HIP graph launch
per ensemble step over uneven grids, checked run by run against the same
runs stepped alone (coral_amd_step), through the C ABI."""
import numpy as np
import pytest

from coral_amd import ic

pytestmark = pytest.mark.gpu

# uneven grids: NX = NY in {64, 96, 128}, NZ in {33, 49, 65}; Ra and Pr per run
RUNS = [(64, 33, 1.3e5, 1.0), (96, 49, 2.2e6, 0.7), (128, 65, 4.0e5, 7.0),
        (64, 65, 8.0e6, 1.0), (128, 33, 1.1e6, 0.7), (96, 33, 3.0e5, 7.0)]


@pytest.fixture(scope="module")
def mods(native_lib):
    from coral_amd import ensemble, solver
    return solver, ensemble


def make_runs(solver, ensemble):
    out = []
    for i, (NX, NZ, Ra, Pr) in enumerate(RUNS):
        r = ensemble.rbc_recipe(Ra, Pr)
        dt = ensemble.run_dt(Ra, Pr)
        k0, z0 = ic.rbc_initial_state(r, NX, NX, NZ, 1e-4, 8, 6, 40 + i)
        g = solver.Solver(r, NX, NX, NZ, 2 * np.pi, 2 * np.pi)
        g.set_state(k0, z0)
        g.factorize_operators(dt)
        out.append((g, dt))
    return out


def states(runs):
    return [g.get_state() for g, _ in runs]


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_ensemble_matches_solo_runs(mods):
    solver, ensemble = mods
    solo = make_runs(solver, ensemble)
    for g, dt in solo:
        g.use_graphs(True)
        g.march_forward(dt, 200)
    ref = states(solo)
    for g, _ in solo:
        g.close()
    for k, z in ref:
        assert all(np.all(np.isfinite(a)) for a in k + z)

    # default launch shapes over three branches: bit for bit the solo steps,
    # also when one run is stepped on its own in between (the ensemble's
    # graphs are recaptured for its swapped buffers)
    runs = make_runs(solver, ensemble)
    with ensemble.Ensemble([g for g, _ in runs]) as ens:
        ens.set_schedule(3)
        ens.step(120)
        g0, dt0 = runs[0]
        g0.march_forward(dt0, 1)
        ens.step(80)
        got = states(runs)
    print("ensemble (3 branches) vs solo after 200 steps: bit-exact check", flush=True)
    for i, ((k, z), (kr, zr)) in enumerate(zip(got, ref)):
        if i == 0:
            continue   # one step ahead, compared below
        for a, b in zip(k + z, kr + zr):
            assert np.array_equal(a, b), f"run {i} differs"
    # run 0 took 201 steps: one more solo step from the reference state
    r0 = solver.Solver(runs[0][0].recipe, RUNS[0][0], RUNS[0][0], RUNS[0][1], 2 * np.pi, 2 * np.pi)
    r0.set_state(*ref[0])
    r0.march_forward(runs[0][1], 1)
    for a, b in zip(sum(got[0], []), sum(r0.get_state(), [])):
        assert np.array_equal(a, b)
    r0.close()
    for g, _ in runs:
        g.close()

    # tuned launch shapes and schedule: the tuner restores every state, and
    # 200 tuned ensemble steps agree with the solo runs to rounding
    runs = make_runs(solver, ensemble)
    with ensemble.Ensemble([g for g, _ in runs]) as ens:
        nb, ms = ens.tune(4)
        print("tuned:", nb, "branches,", ms, "ms per ensemble step;",
              [g.launch_shape() for g, _ in runs], flush=True)
        assert nb != 0 and -len(RUNS) <= nb <= len(RUNS) and ms > 0   # negative: lockstep
        ens.step(200)
        got = states(runs)
    worst = 0.0
    for (k, z), (kr, zr) in zip(got, ref):
        for a, b in zip(k + z, kr + zr):
            worst = max(worst, rel(a, b))
    print("tuned ensemble vs solo after 200 steps: worst rel L2", worst, flush=True)
    assert worst <= 1e-12
    for g, _ in runs:
        g.close()


def test_lockstep_ensemble_matches_solo_runs(mods):
    """The lockstep schedule (set_schedule(-nb), 0 for one branch): one graph
    per ensemble step in which every run's NL terms and right-hand sides of a
    stage (over nb branches) are followed by the stage's kxky solves of all
    runs as one batched launch (k_solve16_batch): bit for bit the solo steps
    after 200 steps, also when one run is stepped on its own in between and
    the schedule changes."""
    solver, ensemble = mods
    solo = make_runs(solver, ensemble)
    for g, dt in solo:
        g.use_graphs(True)
        g.march_forward(dt, 200)
    ref = states(solo)
    for g, _ in solo:
        g.close()
    runs = make_runs(solver, ensemble)
    with ensemble.Ensemble([g for g, _ in runs]) as ens:
        ens.set_schedule(-3)          # stage fronts over three branches
        ens.step(120)
        g0, dt0 = runs[0]
        g0.march_forward(dt0, 1)
        ens.set_schedule(0)           # all on one stream
        ens.step(80)
        got = states(runs)
    print("lockstep ensemble vs solo after 200 steps: bit-exact check", flush=True)
    for i, ((k, z), (kr, zr)) in enumerate(zip(got, ref)):
        if i == 0:
            continue
        for a, b in zip(k + z, kr + zr):
            assert np.array_equal(a, b), f"run {i} differs"
    r0 = solver.Solver(runs[0][0].recipe, RUNS[0][0], RUNS[0][0], RUNS[0][1], 2 * np.pi, 2 * np.pi)
    r0.set_state(*ref[0])
    r0.march_forward(runs[0][1], 1)
    for a, b in zip(sum(got[0], []), sum(r0.get_state(), [])):
        assert np.array_equal(a, b)
    r0.close()
    for g, _ in runs:
        g.close()
