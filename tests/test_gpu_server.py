"""The serving loop: projection requests admitted by memory budget and
advanced together (continuous batching).

Each request's result is bit-identical alone and in a mix with staggered
arrivals, matches the CPU oracle's projection (b = -div u, MG-PCG, u -= dt
grad p) to f32 rounding, and is divergence-free to the solve tolerance.
The budget test checks that admission waits for memory and that every
request still completes.
"""
import numpy as np
import pytest

import panopaea_amd as pa
from oracle import oracle as o

pytestmark = pytest.mark.gpu

TOL = 1e-6
REQS = [
    (dict(dims=(128, 128), dt=0.5 / 128), 1),
    (dict(dims=(32, 32, 32), dt=0.5 / 32, bc=(0, 0, 0, 0, 0, 1)), 2),
    (dict(dims=(192, 64), dt=0.5 / 192), 3),
    (dict(dims=(24, 40, 32), dt=0.5 / 40), 4),
    (dict(dims=(256, 256), dt=0.5 / 256, bc=(0, 0, 0, 1, 0, 0)), 5),
    (dict(dims=(64, 64), dt=0.5 / 64), 6),
]


def modes(g):
    return (16, 8.0) if len(g["dims"]) == 2 else (32, 16.0)


def submit(S, k):
    g, seed = REQS[k]
    nm, km = modes(g)
    return S.submit(g, seed=seed, n_modes=nm, k_max=km)


def run_alone(k):
    S = pa.Server(hbm_budget=1 << 30, max_active=16)
    rid = submit(S, k)
    S.run_until_idle()
    out = S.fetch(rid), S.status(rid)
    S.close()
    return out


def run_mix():
    S = pa.Server(hbm_budget=1 << 30, max_active=16)
    ids = {}
    for k in range(3):
        ids[k] = submit(S, k)
    S.tick()
    S.tick()
    for k in range(3, len(REQS)):  # arrive while the first three iterate
        ids[k] = submit(S, k)
        S.tick()
    S.run_until_idle()
    out = {k: (S.fetch(i), S.status(i)) for k, i in ids.items()}
    S.close()
    return out


def test_gpu_server_alone_equals_mix(ctx):
    mix = run_mix()
    for k in range(len(REQS)):
        (f1, p1), s1 = run_alone(k)
        (f2, p2), s2 = mix[k]
        assert s1["state"] == s2["state"] == pa.Server.DONE
        assert s1["converged"] == 1 and s1["iterations"] == s2["iterations"]
        assert np.array_equal(f1, f2) and np.array_equal(p1, p2), REQS[k][0]


def test_gpu_server_matches_oracle_projection(ctx):
    S = pa.Server(hbm_budget=1 << 30, max_active=16)
    ids = []
    hosts = []
    for k, (g, seed) in enumerate(REQS[:4]):
        grid = (pa.Grid2d if len(g["dims"]) == 2 else pa.Grid3d)(g["dims"], np.float32, bc=g.get("bc"))
        nm, km = modes(g)
        faces = pa.gen_bandlimited_faces(grid, seed, nm, km)
        hosts.append(faces)
        ids.append(S.submit(g, faces=faces))  # host upload path
    S.run_until_idle()
    for (g, _), faces, rid in zip(REQS[:4], hosts, ids):
        f, p = S.fetch(rid)
        st = S.status(rid)
        P = o.params(g["dims"], bc=g.get("bc", (0,) * 6), dt=g["dt"], precond="mg", levels=4, tol=TOL,
                     max_iterations=500)
        b = o.neg_div(P, faces)
        x, info, _ = o.solve(P, b)
        u = o.grad_sub(P, faces, x)
        assert info.iterations == st["iterations"]
        scale = np.abs(faces).max()
        assert np.abs(f - u).max() <= 1e-5 * scale, g
        # divergence left: the solve's residual plus the f32 rounding of
        # u - dt grad p, which the oracle's f32 projection shows as well
        div = np.abs(o.neg_div(P, f)).max()
        div_ref = np.abs(o.neg_div(P, u)).max()
        assert div <= 2 * max(div_ref, TOL * np.abs(b).max()), (g, div, div_ref)


def test_gpu_server_budget_admission(ctx):
    # room for about two 256^2 requests at a time: the rest wait in the queue
    # until completed requests are fetched and released
    g = dict(dims=(256, 256), dt=0.5 / 256)
    budget = 5 << 20
    S = pa.Server(hbm_budget=budget, max_active=16)
    ids = [S.submit(g, seed=s) for s in range(6)]
    t = S.tick()
    assert 0 < t["active"] < 6 and t["queued"] > 0
    assert t["bytes_used"] <= budget
    results = {}

    def on_done(rid):
        st = S.status(rid)
        assert st["state"] == pa.Server.DONE and st["converged"] == 1
        results[rid] = S.fetch(rid)
        S.release(rid)

    S.run_until_idle(on_done)
    assert sorted(results) == sorted(ids)
    assert S.tick()["bytes_used"] == 0
    S.close()
    alone = run_alone_grid(g, 3)
    f, p = results[ids[3]]
    assert np.array_equal(f, alone[0]) and np.array_equal(p, alone[1])


def run_alone_grid(g, seed):
    S = pa.Server(hbm_budget=1 << 30, max_active=16)
    rid = S.submit(g, seed=seed)
    S.run_until_idle()
    out = S.fetch(rid)
    S.close()
    return out
