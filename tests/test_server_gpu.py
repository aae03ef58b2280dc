"""GPU tests of the live request server (svo_server_*): requests submitted
at any time, from several host threads, run as their own launches or
together in batched launches (or, with persistent_launch, folded into one
running persistent launch); each result must equal the same request rendered
alone by the plain kernel, bit for bit, and the oracle on sampled pixels.
"""
import os
import threading
import time

import numpy as np
import pytest

import oracle_py as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trees(svo, golden):
    return [svo.VoxelOctree(os.path.join(golden, "dragon.oct")),
            svo.VoxelOctree(os.path.join(golden, "build_sdf128.oct"))]


def _requests(trees, n, seed, wmax=400, hmax=300):
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(n):
        ti = int(rng.integers(0, len(trees)))
        w, h = int(rng.integers(1, wmax)), int(rng.integers(1, hmax))
        cam = trees[ti].orbit_camera(float(rng.uniform(-60, 60)), float(rng.uniform(0, 360)), 1.0)
        reqs.append((ti, cam, w, h))
    return reqs


def _check(trees, req, out, oracle_every=0):
    ti, cam, w, h = req
    ref = trees[ti].render_primary(cam, w, h)
    assert np.array_equal(out["normal"], ref["normal"])
    assert np.array_equal(out["t"].view(np.uint32), ref["t"].view(np.uint32))
    assert np.array_equal(out["leaf"], ref["leaf"])
    if oracle_every:
        sample = np.arange(0, w * h, oracle_every)
        d = orc.primary_dirs(np.frombuffer(cam.tform, np.float32), w, h)[sample]
        o = orc.raymarch(trees[ti].words(), np.frombuffer(cam.pos, np.float32)[None, :], d, 0.0)
        assert np.array_equal(out["normal"][sample], o["normal"])
        assert np.array_equal(out["t"][sample].view(np.uint32), o["t"].view(np.uint32))
        assert np.array_equal(out["leaf"][sample], o["leaf"])


@pytest.mark.parametrize("host_outputs,persistent", [(False, False), (True, False), (False, True), (True, True)])
def test_results_equal_alone_and_in_the_mix(svo, trees, host_outputs, persistent):
    """240 requests submitted from 4 host threads at once (sizes 1x1 to
    400x300 over two octrees): every result equals the request rendered alone
    and the oracle on every 13th pixel; by default no persistent launch runs
    and the requests that queue up run in batched launches, with
    persistent_launch they go to the persistent launch."""
    reqs = _requests(trees, 240, seed=5)
    print(f"[gpu] live server, 4 submitting threads, host_outputs={host_outputs}, persistent={persistent}",
          flush=True)
    srv = svo.svo.Server(trees, output_budget=1 << 30, host_outputs=host_outputs, persistent_launch=persistent)
    tickets = [None] * len(reqs)
    errors = []

    def work(k):
        try:
            for i in range(k, len(reqs), 4):
                ti, cam, w, h = reqs[i]
                tickets[i] = srv.submit(ti, cam, w, h)
        except Exception as e:  # reported below
            errors.append(repr(e))

    th = [threading.Thread(target=work, args=(k,)) for k in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    for i, tk in enumerate(tickets):
        lat = srv.wait(tk, timeout=60.0)
        assert lat > 0
        _, _, w, h = reqs[i]
        _check(trees, reqs[i], srv.copy_outputs(tk, w, h), oracle_every=13 if i % 8 == 0 else 0)
        srv.release(tk)
    st = srv.stats()
    assert st["published"] == len(reqs) and st["in_flight"] == 0
    if persistent:
        assert st["launches"] >= 1 and st["batched"] == 0
    else:
        assert st["launches"] == 0 and st["batched"] > 0 and st["direct"] + st["batched"] == len(reqs)
    srv.close()


@pytest.mark.parametrize("persistent", [False, True])
def test_admission_by_output_footprint(svo, trees, persistent):
    """An arena that holds about two 320x240 requests: 30 such requests wait
    for memory, are admitted as finished ones are released and all complete
    correctly; a request larger than the arena is refused."""
    budget = 2 * 320 * 240 * 16 + 4096
    srv = svo.svo.Server(trees, output_budget=budget, persistent_launch=persistent)
    reqs = _requests(trees, 30, seed=11, wmax=2, hmax=2)
    reqs = [(ti, cam, 320, 240) for ti, cam, _, _ in reqs]
    print(f"[gpu] live server admission by footprint, persistent={persistent}", flush=True)
    tickets = [srv.submit(*r) for r in reqs]
    assert srv.stats()["waiting"] > 0
    for r, tk in zip(reqs, tickets):
        srv.wait(tk, timeout=60.0)
        _check(trees, r, srv.copy_outputs(tk, 320, 240))
        srv.release(tk)
    with pytest.raises(svo.SvoError) as e:
        srv.submit(0, reqs[0][1], 1920, 1080)
    assert e.value.code == svo.svo.ERR_LIMIT
    with pytest.raises(svo.SvoError) as e:
        srv.submit(7, reqs[0][1], 8, 8)
    assert e.value.code == svo.svo.ERR_INVALID
    srv.close()


def test_launch_ends_when_idle_and_restarts(svo, trees):
    """With persistent_launch, requests beyond the four that run as their own
    launches go to the persistent launch; it ends after idle_stop_us with
    nothing in flight, and the next burst starts a new one; every result is
    traced correctly."""
    srv = svo.svo.Server(trees, idle_stop_us=1000, persistent_launch=True)
    reqs = _requests(trees, 16, seed=23)
    print("[gpu] live server idle stop / restart", flush=True)

    def burst(rs):
        tks = [srv.submit(*r) for r in rs]
        for r, tk in zip(rs, tks):
            srv.wait(tk, timeout=60.0)
            _check(trees, r, srv.copy_outputs(tk, r[2], r[3]))
            srv.release(tk)

    burst(reqs[:8])
    deadline = time.time() + 5.0
    while srv.stats()["live"] and time.time() < deadline:
        time.sleep(0.01)
    assert srv.stats()["live"] == 0
    first = srv.stats()["launches"]
    assert first >= 1
    burst(reqs[8:])
    assert srv.stats()["launches"] >= first + 1
    srv.close()


@pytest.mark.parametrize("persistent_only", [False, True])
def test_lone_requests_run_as_their_own_launch(svo, trees, persistent_only):
    """A request submitted with nothing in flight runs as an ordinary launch
    (stats: direct) unless persistent_only; either way its result equals the
    request rendered alone and the oracle on every 7th pixel."""
    srv = svo.svo.Server(trees, idle_stop_us=1000, persistent_only=persistent_only)
    reqs = _requests(trees, 4, seed=41)
    print(f"[gpu] live server lone requests, persistent_only={persistent_only}", flush=True)
    for r in reqs:
        tk = srv.submit(*r)
        srv.wait(tk, timeout=60.0)
        _check(trees, r, srv.copy_outputs(tk, r[2], r[3]), oracle_every=7)
        srv.release(tk)
    st = srv.stats()
    assert st["published"] == len(reqs) and st["in_flight"] == 0
    if persistent_only:
        assert st["direct"] == 0 and st["launches"] >= 1
    else:
        assert st["direct"] == len(reqs) and st["launches"] == 0
    srv.close()


def test_schedule_keeps_results_until_released(svo, trees):
    """run_schedule with a Poisson schedule (3000 req/s, 120 requests) and
    kept outputs: latencies are positive, every kept result is correct."""
    reqs = _requests(trees, 120, seed=31, wmax=256, hmax=192)
    rng = np.random.default_rng(3)
    arr = np.cumsum(rng.exponential(1.0 / 3000.0, len(reqs)))
    arr -= arr[0]
    srv = svo.svo.Server(trees, output_budget=1 << 28)
    print("[gpu] live server schedule, 120 requests", flush=True)
    lat, mk, tickets = srv.run_schedule(reqs, arr, keep=True)
    assert np.all(lat > 0) and mk > 0
    for r, tk in zip(reqs, tickets):
        _check(trees, r, srv.copy_outputs(int(tk), r[2], r[3]))
        srv.release(int(tk))
    srv.close()
