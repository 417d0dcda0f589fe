"""GPU parity of the serving engine (msl_server_*, msl_serve_plain_loop).

Bar:
  * each job's final state is bit-identical to the oracle run on that job
    alone (blur: mean blur iterated; heat: 7-point heat iterated; map+fold:
    x*mul+add in uint32);
  * float sums within FOLD_REL_TOL of the FP64 sum of the final state, u32
    sums exact;
  * every result (final state and sum, bit for bit) is the same whether the
    job is served alone or in a mix, with any arrival pattern and admission
    budget;
  * the plain one-launch-per-step loop gives the same final sums within
    FOLD_REL_TOL (different fold partition) and identical u32 sums.
"""
import math
import sys

import numpy as np
import pytest

from muesli4_amd import hip
from muesli4_amd.hip import MslError
from muesli4_amd.serve import Server, job_bytes, plain_loop
from muesli4_amd.workloads import JOB_BLUR2D, JOB_HEAT3D, JOB_MAPFOLD_U32, serve_trace
from oracle import refcpu as R

pytestmark = pytest.mark.gpu

FOLD_REL_TOL = 1e-6


def note(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _init():
    hip.init(0)
    yield
    hip.synchronize()


def small_trace(n=30, seed=11, rate=0.0):
    jobs, arr = serve_trace(n, seed, rate, blur_side=(8, 160), blur_steps=(1, 9), heat_side=(6, 40),
                            heat_steps=(1, 6), map_len=(1e2, 3e5))
    for k, d in enumerate(jobs):
        if d["kind"] != JOB_MAPFOLD_U32:
            d["nv"] = 0.25 if k % 4 == 0 else 0.0
    return jobs, arr


def oracle_job(d):
    if d["kind"] == JOB_BLUR2D:
        x = R.fill_uniform_f32(d["d0"] * d["d1"], d["seed"]).reshape(d["d0"], d["d1"])
        for _ in range(d["steps"]):
            x = R.stencil2d_mean_f32(x, d["r"], d.get("nv", 0.0))
        return x
    if d["kind"] == JOB_HEAT3D:
        x = R.fill_uniform_f32(d["d0"] * d["d1"] * d["d2"], d["seed"], R.SCALE_UNIT).reshape(d["d2"], d["d1"], d["d0"])
        for _ in range(d["steps"]):
            x = R.heat7_f32(x, d.get("nv", 0.0))
        return x
    x = R.fill_u32(d["n"], d["seed"])
    return (x.astype(np.uint64) * d["mul"] + d["add"]).astype(np.uint32)


def check_job(d, res):
    want = oracle_job(d)
    np.testing.assert_array_equal(res["final"], want, err_msg=str(d))
    if d["kind"] == JOB_MAPFOLD_U32:
        assert res["sum_u32"] == int(want.astype(np.uint64).sum() & 0xFFFFFFFF)
    else:
        ref = math.fsum(want.astype(np.float64).ravel())
        assert abs(res["sum"] - ref) <= FOLD_REL_TOL * max(abs(ref), 1e-30), (d, res["sum"], ref)


def serve_all(jobs, budget=1 << 30, max_active=64, steps_between=0):
    """Submit every job (keep_final), with optional rounds between
    submissions so later jobs join a running batch; then step and poll until
    every job is done (polling releases a job's memory)."""
    srv = Server(budget, max_active)
    ids = []
    out = {}

    def poll_all():
        for i, d in zip(ids, jobs):
            if i not in out:
                r = srv.poll(i, d)
                if r is not None:
                    out[i] = r

    for d in jobs:
        ids.append(srv.submit(d, keep_final=True))
        for _ in range(steps_between):
            srv.step()
            poll_all()
    for _ in range(100000):
        poll_all()
        if len(out) == len(ids):
            break
        srv.step()
    assert len(out) == len(ids)
    info = srv.info()
    srv.close()
    return [out[i] for i in ids], info


def test_server_results_match_oracle():
    note("server mix vs oracle")
    jobs, _ = small_trace()
    out, info = serve_all(jobs)
    assert info["done_jobs"] == len(jobs)
    for d, res in zip(jobs, out):
        check_job(d, res)


def test_alone_equals_mix():
    note("server alone vs mix")
    jobs, _ = small_trace(n=16, seed=12)
    mix, _ = serve_all(jobs, steps_between=2)
    for d, m in zip(jobs, mix):
        alone, _ = serve_all([d])
        np.testing.assert_array_equal(alone[0]["final"], m["final"])
        assert alone[0]["sum"] == m["sum"] and alone[0]["sum_u32"] == m["sum_u32"]


def test_light_and_heavy_load_tilings_agree():
    """A job served alone runs at light load (short segments, several steps
    per round); in a mix whose items fill the GPU it runs with the default
    segments, one step per round.  Final state and sum are bit-identical
    either way, and match the oracle."""
    note("server light vs heavy load tiling")
    blur = dict(kind=JOB_BLUR2D, d0=1000, d1=1100, r=3, steps=12, seed=21, nv=0.5)
    heavy = [dict(kind=JOB_HEAT3D, d0=256, d1=256, d2=256, steps=3, seed=22 + k) for k in range(2)]
    alone, _ = serve_all([blur])
    mix, _ = serve_all(heavy + [blur], budget=2 << 30)
    np.testing.assert_array_equal(alone[0]["final"], mix[2]["final"])
    assert alone[0]["sum"] == mix[2]["sum"]
    check_job(blur, alone[0])
    for d, res in zip(heavy, mix):
        check_job(d, res)


def test_tight_budget_admission():
    """A budget that holds only the largest job: jobs wait for memory, the
    results do not change."""
    note("server tight budget")
    jobs, _ = small_trace(n=20, seed=13)
    budget = max(job_bytes(d) for d in jobs)
    tight, info = serve_all(jobs, budget=budget, max_active=64)
    assert info["peak_bytes"] <= budget
    roomy, _ = serve_all(jobs)
    for a, b in zip(tight, roomy):
        np.testing.assert_array_equal(a["final"], b["final"])
        assert a["sum"] == b["sum"] and a["sum_u32"] == b["sum_u32"]


def test_slot_limit():
    note("server one slot")
    jobs, _ = small_trace(n=8, seed=14)
    one, info = serve_all(jobs, max_active=1)
    many, _ = serve_all(jobs)
    for a, b in zip(one, many):
        np.testing.assert_array_equal(a["final"], b["final"])
        assert a["sum"] == b["sum"]


def test_run_trace_and_plain_loop_agree():
    note("server trace vs plain loop")
    jobs, arr = small_trace(n=40, seed=15, rate=4000.0)
    srv = Server(1 << 30, 64)
    res = srv.run_trace(jobs, arr)
    srv.close()
    base = plain_loop(jobs, arr, 1 << 30)
    mix, _ = serve_all(jobs)
    for d, r, b, m in zip(jobs, res, base, mix):
        assert r["t_arrival"] <= r["t_admit"] <= r["t_done"]
        if d["kind"] == JOB_MAPFOLD_U32:
            assert r["sum_u32"] == b["sum_u32"] == m["sum_u32"]
        else:
            assert r["sum"] == m["sum"]  # same fold partition (fixed chunks of the final state)
            assert abs(r["sum"] - b["sum"]) <= 2 * FOLD_REL_TOL * max(abs(b["sum"]), 1e-30)


def test_server_rejects_bad_jobs():
    srv = Server(1 << 20, 4)
    with pytest.raises(MslError):
        srv.submit(dict(kind=JOB_BLUR2D, d0=16, d1=16, r=11, steps=1, seed=1))
    with pytest.raises(MslError):  # larger than the budget
        srv.submit(dict(kind=JOB_HEAT3D, d0=128, d1=128, d2=128, steps=1, seed=1))
    with pytest.raises(MslError):
        srv.submit(dict(kind=9, seed=1))
    jid = srv.submit(dict(kind=JOB_MAPFOLD_U32, n=1000, mul=3, add=1, seed=2))
    assert srv.poll(jid) is None  # not run yet
    srv.drain()
    assert srv.poll(jid) is not None
    srv.close()


@pytest.mark.slow
def test_full_size_trace_server_vs_plain_loop():
    """The serving configuration (sides up to 2048 / 256, up to 500 steps,
    lengths up to 1e8) at saturation: every u32 sum identical, float sums
    within tolerance of the plain loop."""
    note("server full-size trace")
    jobs, arr = serve_trace(60, 2025, 0.0)
    srv = Server(32 << 30, 1024)
    res = srv.run_trace(jobs, arr)
    srv.close()
    base = plain_loop(jobs, arr, 32 << 30)
    for d, r, b in zip(jobs, res, base):
        if d["kind"] == JOB_MAPFOLD_U32:
            assert r["sum_u32"] == b["sum_u32"]
        else:
            assert abs(r["sum"] - b["sum"]) <= 2 * FOLD_REL_TOL * abs(b["sum"]), (d, r["sum"], b["sum"])
