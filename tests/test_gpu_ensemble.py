"""Ensemble serving: each box's result in the mix equals the result alone
(bit for bit: the kernels are deterministic and boxes share no state), HBM
budget admission, and a stop in one box leaves the others untouched."""
import numpy as np
import pytest

from mhit36_amd import Ensemble, Solver, default_params, MhitError, MHIT_ERR_CFL
from mhit36_amd._lib import MHIT_ERR_ARG
from mhit36_amd import ic

pytestmark = pytest.mark.gpu

BOXES = [
    dict(nx=16, dt=1e-3, mu=0.01, phiflag=0, precision=0),
    dict(nx=32, dt=5e-4, mu=0.01, phiflag=1, radius=1.0, sigma=0.1, precision=0),
    dict(nx=24, dt=1e-3, mu=0.02, phiflag=0, f1=0.5, precision=1),
    dict(nx=48, dt=5e-4, mu=0.01, phiflag=1, radius=0.8, sigma=0.1, gamma_factor=1.2, precision=1),
    dict(nx=64, dt=5e-4, mu=0.005, phiflag=1, radius=0.9, sigma=0.05, precision=0),
]
STEPS = [7, 12, 5, 9, 6]


def alone(p, nsteps):
    with Solver(**p) as s:
        s.init_fields()
        s.step(nsteps)
        return {f: s.get_field(f) for f in ((1, 2, 3, 4, 5) if p["phiflag"] else (1, 2, 3, 4))}


@pytest.mark.parametrize("graphs,gang", [(True, 1), (False, 1), (True, 8)])
def test_same_alone_and_in_mix(graphs, gang):
    ref = [alone(p, n) for p, n in zip(BOXES, STEPS)]
    with Ensemble(use_graphs=graphs, inflight=3, gang_size=gang) as ens:
        ids = [ens.submit(default_params(**p), nsteps=n) for p, n in zip(BOXES, STEPS)]
        ens.wait_idle()
        res = {r["id"]: r for r in ens.results()}
        for q, rid in enumerate(ids):
            assert res[rid]["status"] == 0 and res[rid]["steps_done"] == STEPS[q]
            for f, a in ref[q].items():
                assert np.array_equal(ens.get_field(rid, f), a), (q, f)
        lat = ens.latency()
        # one latency per step after each box's first (gaps between completions)
        assert lat["steps"] == sum(STEPS) - len(STEPS) and lat["p99_ms"] >= lat["p50_ms"] > 0


def test_budget_admission_serialises():
    """A budget that holds one 64^3 box at a time admits the next box only
    after the previous one finished; results are unchanged."""
    p = dict(nx=64, dt=5e-4, mu=0.005, phiflag=0, precision=0)
    ref = alone(p, 4)
    with Ensemble(hbm_budget_bytes=30 << 20, max_active=8) as ens:      # one 64^3 FP64 box is ~23 MB
        ids = [ens.submit(default_params(**p), nsteps=4) for _ in range(3)]
        ens.wait_idle()
        res = sorted(ens.results(), key=lambda r: r["id"])
        for a, b in zip(res, res[1:]):
            assert b["admit_ms"] >= a["finish_ms"]
        for rid in ids:
            assert np.array_equal(ens.get_field(rid, 1), ref[1])


@pytest.mark.parametrize("bad_nx", [16, 32])      # 32: the stopping box shares the good box's gang
def test_halted_box_does_not_disturb_others(bad_nx):
    good = dict(nx=32, dt=1e-3, mu=0.01, phiflag=0, precision=0)
    bad = dict(nx=bad_nx, dt=3.0, mu=0.0, phiflag=0, precision=0)
    ref = alone(good, 10)
    with Ensemble() as ens:
        g = ens.submit(default_params(**good), nsteps=10)
        b = ens.submit(default_params(**bad), nsteps=10)
        ens.wait_idle()
        res = {r["id"]: r for r in ens.results()}
        assert res[b]["status"] == MHIT_ERR_CFL and res[b]["steps_done"] == 1
        assert res[g]["status"] == 0
        assert np.array_equal(ens.get_field(g, 1), ref[1])


def test_submitted_fields_are_used():
    n = 32
    u, v, w = ic.random_velocity(n, seed=3)
    p = dict(nx=n, dt=1e-3, mu=0.01, phiflag=0, inflow=1, precision=0)
    with Solver(**p) as s:
        for f, a in ((1, u), (2, v), (3, w)):
            s.set_field(f, a)
        s.step(5)
        ref = s.get_field(1)
    with Ensemble() as ens:
        rid = ens.submit(default_params(**p), nsteps=5, u=u, v=v, w=w)
        ens.wait_idle()
        assert np.array_equal(ens.get_field(rid, 1), ref)


def test_submit_rejects_incomplete_initial_fields():
    """u, v, w come all three or not at all; inflow = 1 needs them and
    inphi = 1 (with the phase field on) needs phi: a request is never run on
    an initial condition other than the one asked for."""
    n = 16
    u, v, w = ic.random_velocity(n, seed=4)
    with Ensemble() as ens:
        for kw, fields in ((dict(phiflag=0), dict(u=u)),
                           (dict(phiflag=0), dict(u=u, v=v)),
                           (dict(phiflag=0, inflow=1), {}),
                           (dict(phiflag=1, inphi=1), {})):
            with pytest.raises(MhitError) as e:
                ens.submit(default_params(nx=n, dt=1e-3, mu=0.01, **kw), nsteps=1, **fields)
            assert e.value.code == MHIT_ERR_ARG, (kw, sorted(fields))
        rid = ens.submit(default_params(nx=n, dt=1e-3, mu=0.01, phiflag=0, inflow=1), nsteps=1, u=u, v=v, w=w)
        ens.wait_idle()
        assert {r["id"]: r for r in ens.results()}[rid]["status"] == 0


def test_close_with_boxes_in_flight():
    """Destroying the ensemble while boxes are admitted, stepping and queued
    returns cleanly (I/O threads joined, streams drained)."""
    p = dict(nx=64, dt=5e-4, mu=0.005, phiflag=1, radius=0.8, sigma=0.1, precision=1)
    ens = Ensemble(max_active=2, inflight=3)
    for _ in range(5):
        ens.submit(default_params(**p), nsteps=500)
    for _ in range(20):
        ens.progress()
    ens.close()


def test_results_arrive_while_others_run():
    """Continuous batching: a short box finishes, and a box submitted later is
    admitted, while a long box keeps running."""
    long_p = dict(nx=96, dt=5e-4, mu=0.005, phiflag=0, precision=0)
    short_p = dict(nx=32, dt=1e-3, mu=0.01, phiflag=0, precision=0)
    with Ensemble(inflight=2) as ens:
        lid = ens.submit(default_params(**long_p), nsteps=400)
        sid = ens.submit(default_params(**short_p), nsteps=5)
        done = {}
        late = None
        while True:
            act, q = ens.progress()
            for r in ens.results():
                done[r["id"]] = r
            if sid in done and late is None:
                assert lid not in done            # the long box is still running
                late = ens.submit(default_params(**short_p), nsteps=5)
            if act == 0 and q == 0 and late is not None:
                break
        for r in ens.results():
            done[r["id"]] = r
        assert done[late]["admit_ms"] < done[lid]["finish_ms"]
        assert np.array_equal(ens.get_field(late, 1), ens.get_field(sid, 1))


def test_gangs_take_and_release_boxes():
    """Continuous batching inside gangs: twelve boxes of one class with
    different step counts, submitted in three waves while earlier boxes run,
    through gangs of four (members leave when done, later boxes join between
    steps).  Every box equals the same box alone, and every step after each
    box's first has a latency."""
    _, dx = ic.grid(32)
    ps, steps = [], []
    for q in range(12):
        p = dict(nx=32, dt=(0.04 + 0.005 * q) * dx, mu=0.004 + 0.002 * q, phiflag=1, radius=0.5,
                 sigma=0.05, gamma_factor=1.2, precision=1, f1=0.5 * (q % 3))
        ps.append(p)
        steps.append(3 + (7 * q) % 11)
    ref = [alone(p, n) for p, n in zip(ps, steps)]
    with Ensemble(inflight=2, gang_size=4) as ens:
        ids = []
        for wave in range(3):
            for q in range(4 * wave, 4 * wave + 4):
                ids.append(ens.submit(default_params(**ps[q]), nsteps=steps[q]))
            for _ in range(3):
                ens.progress()
        ens.wait_idle()
        res = {r["id"]: r for r in ens.results()}
        for q, rid in enumerate(ids):
            assert res[rid]["status"] == 0 and res[rid]["steps_done"] == steps[q]
            for f, a in ref[q].items():
                assert np.array_equal(ens.get_field(rid, f), a), (q, f)
        assert ens.latency()["steps"] == sum(steps) - len(steps)


@pytest.mark.parametrize("graphs", [True, False])
def test_gang_graph_replays_match_alone(graphs):
    """Long gang memberships: after a few steps a steady gang replays captured
    graphs (one per phi parity); the membership changes as boxes finish (20,
    30 and 40 steps), so graphs are captured again for the smaller gangs.
    Every box still equals the same box alone."""
    _, dx = ic.grid(64)
    ps = [dict(nx=64, dt=(0.04 + 0.01 * q) * dx, mu=0.003 + 0.002 * q, phiflag=1, radius=0.6 + 0.1 * q,
               sigma=0.05, gamma_factor=1.2, precision=1) for q in range(4)]
    steps = [20, 30, 40, 40]
    ref = [alone(p, n) for p, n in zip(ps, steps)]
    with Ensemble(inflight=2, gang_size=4, use_graphs=graphs) as ens:
        ids = [ens.submit(default_params(**p), nsteps=n) for p, n in zip(ps, steps)]
        ens.wait_idle()
        res = {r["id"]: r for r in ens.results()}
        for q, rid in enumerate(ids):
            assert res[rid]["status"] == 0 and res[rid]["steps_done"] == steps[q]
            for f, a in ref[q].items():
                assert np.array_equal(ens.get_field(rid, f), a), (q, f)
