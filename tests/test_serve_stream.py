"""The serving stream generator: seeded and reproducible, and every request
inside the ranges of the serving configuration (N from the size list,
nu in [0.002, 0.02], about half with 8-64 drops of radius 0.15-0.3, step
counts in range, mixed precision, Poisson arrivals)."""
import numpy as np

from mhit36_amd import serve


def stream(seed, count=12):
    return serve.serving_stream(count=count, seed=seed, steps=(200, 2000), sizes=(64,))


def test_reproducible():
    a, b = stream(5), stream(5)
    for x, y in zip(a, b):
        assert x["arrival"] == y["arrival"] and x["params"] == y["params"] and x["nsteps"] == y["nsteps"]
        assert (x["phi"] is None) == (y["phi"] is None)
        if x["phi"] is not None:
            assert np.array_equal(x["phi"], y["phi"])
    assert [r["params"] for r in serve.serving_stream(count=3, seed=6, sizes=(64,))] != [r["params"] for r in a[:3]]


def test_ranges():
    reqs = stream(11, count=24)
    t = [r["arrival"] for r in reqs]
    assert all(b > a for a, b in zip(t, t[1:]))
    acdi = 0
    for r in reqs:
        p = r["params"]
        assert p["nx"] == 64 and 0.002 <= p["mu"] <= 0.02 and p["precision"] == 1
        assert 200 <= r["nsteps"] <= 2000
        if p["phiflag"]:
            acdi += 1
            assert 0.15 <= p["radius"] <= 0.3 and r["phi"].shape == (p["nx"],) * 3
            assert 0.0 <= r["phi"].min() and r["phi"].max() <= 1.0
        else:
            assert r["phi"] is None
    assert 4 <= acdi <= 20


def test_cell_steps():
    reqs = serve.serving_stream(count=4, seed=3, sizes=(64,))
    assert serve.cell_steps(reqs) == sum(r["params"]["nx"] ** 3 * r["nsteps"] for r in reqs)
