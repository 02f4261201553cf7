"""Serving with continuous batching: results are the same alone and in the
mix, admission respects the device-memory budget, and late arrivals join
running solves."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no HIP device")
    return t


@pytest.fixture(scope="module")
def S(torch):
    import ravenhip.serve as serve
    return serve


def small_stream(S, arrivals):
    specs = [(2, "advection", 5, 9, 0.4, "f64"), (3, "euler", 6, 3, 0.3, "f32"),
             (2, "euler", 8, 5, 0.25, "f32"), (3, "advection", 4, 4, 0.5, "f64"),
             (2, "euler", 4, 12, 0.35, "f64"), (3, "euler", 8, 2, 0.2, "f64"),
             (2, "advection", 7, 6, 0.15, "f32")]
    return [S.SolveRequest(rid=i, arrival=t, dim=d, law=law, nq=nq, n=n, T=T, dtype=dt, seed=40 + i)
            for i, ((d, law, nq, n, T, dt), t) in enumerate(zip(specs, arrivals))]


def test_results_same_alone_and_in_the_mix(torch, S):
    mixed = small_stream(S, [0.0] * 7)
    # admit the whole burst in one pass: all seven solves step together
    server = S.Server(hbm_budget=8 << 30, admit_slice=60.0)
    server.run(mixed, keep_results=True)
    assert server.peak_active == 7
    alone = small_stream(S, [0.0] * 7)
    S.run_plain(alone, keep_results=True)
    for m, a in zip(mixed, alone):
        print(f"   request {m.rid}: {m.dim}-D {m.law} Nq={m.nq} {m.dtype}: {m.steps} steps",
              flush=True)
        assert m.steps == a.steps > 0
        assert torch.equal(m.result, a.result), f"request {m.rid} differs"


def test_admission_respects_budget_and_late_arrivals_join(torch, S):
    reqs = small_stream(S, [0.0, 0.0, 0.05, 0.05, 0.1, 0.1, 0.1])
    budget = max(r.footprint() for r in reqs) * 2
    server = S.Server(hbm_budget=budget)
    server.run(reqs, keep_results=True)
    assert server.peak_bytes <= budget
    assert all(np.isfinite(r.finished) and r.finished >= r.admitted >= r.arrival for r in reqs)
    alone = small_stream(S, [0.0] * 7)
    S.run_plain(alone, keep_results=True)
    for m, a in zip(reqs, alone):
        assert torch.equal(m.result, a.result), f"request {m.rid} differs"
    summary = S.summarize(reqs, max(r.finished for r in reqs))
    assert summary["p99_time_to_solution_s"] >= summary["p50_time_to_solution_s"] > 0
