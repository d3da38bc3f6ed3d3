"""Solve server on the GPU: results in a mix equal results alone (same bits),
the HBM budget holds, and failure / cancel paths reclaim their slots."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MIX = [((24, 20, 28), 0.1, 150.0), ((9, 13, 7), 0.1, 60.0), ((40, 30, 36), 0.1, 900.0),
       ((16, 16, 16), 0.1, 100.0), ((31, 45, 27), 0.1, 1200.0), ((12, 9, 18), 0.1, 50.0)]


def requests(tol=1e-6, max_steps=600, keep=True):
    from cavityflow_cuda_amd.serving import SolveRequest
    return [SolveRequest(i, s, u, re, tol=tol, max_steps=max_steps, keep_fields=keep)
            for i, (s, u, re) in enumerate(MIX)]


def serve_staggered(srv, reqs):
    """Submit two requests per round: members join while others run."""
    todo = list(reqs)
    while todo or not srv.idle():
        for _ in range(2):
            if todo:
                srv.submit(todo.pop(0))
        srv.round()
    return srv.results


PAIRS = dict(block=256, cells_per_thread=1, waves_per_cu=0, options=(2, 0), pairs=1)


@pytest.mark.parametrize("pairs,graph", [(0, True), (1, True), (0, False)])
@pytest.mark.parametrize("precision", ["f32", "f16s"])
def test_results_same_alone_and_in_the_mix(hip, precision, pairs, graph):
    from cavityflow_cuda_amd.serving import CavityServer, solve_sequential
    print("[case] serving mix", precision, "pairs", pairs, "graph", graph, flush=True)
    # tol 2e-4 lets the small, viscous cavities converge early, the others hit the cap;
    # the solo runs always use the in-place half-step launches; graph=False issues
    # each round's launches directly
    srv = CavityServer(precision, check_every=52, tuning=PAIRS if pairs else None, graph=graph)
    mixed = serve_staggered(srv, requests(tol=2e-4, max_steps=624))
    alone = solve_sequential(requests(tol=2e-4, max_steps=624), precision, check_every=52)
    statuses = set()
    for i in range(len(MIX)):
        a, b = mixed[i], alone[i]
        statuses.add(a.status)
        assert a.status == b.status and a.steps == b.steps, (i, a.status, b.status)
        assert a.status in ("converged", "step_cap"), (i, a.message)
        assert a.fields.tobytes() == b.fields.tobytes(), i
        assert a.totals.tobytes() == b.totals.tobytes(), i
        assert a.residual == b.residual, i      # the batch residual has the same bits
    assert statuses == {"converged", "step_cap"}
    srv.close()


def test_budget_bounds_resident_memory(hip):
    from cavityflow_cuda_amd.serving import CavityServer
    from cavityflow_cuda_amd.solver import device_bytes
    print("[case] serving budget", flush=True)
    reqs = requests(max_steps=200, keep=False)
    biggest = max(device_bytes(r.size) for r in reqs)
    srv = CavityServer("f32", hbm_budget=biggest + device_bytes((16, 16, 16)))
    for r in reqs:
        srv.submit(r)
    most = 0
    while not srv.idle():
        srv.round()
        assert srv.used <= srv.budget
        most = max(most, len(srv.active))
    assert srv.peak_used <= srv.budget and srv.used == 0
    assert all(srv.results[i].status == "step_cap" for i in range(len(reqs)))
    assert most < len(reqs)
    srv.close()


def test_failed_and_cancelled_requests_release_their_slots(hip):
    from cavityflow_cuda_amd.serving import CavityServer, SolveRequest
    print("[case] serving failure paths", flush=True)
    size = (10, 12, 14)
    bad = np.zeros((19, size[2], size[1], size[0]), np.float32)
    bad[:] = np.nan
    srv = CavityServer("f32", check_every=20)
    srv.submit(SolveRequest(0, size, 0.1, 100.0, max_steps=200, initial=bad))
    srv.submit(SolveRequest(1, (20, 20, 20), 0.1, 100.0, max_steps=2000))
    srv.submit(SolveRequest(2, (8, 8, 8), 0.1, 100.0, max_steps=200))
    srv.round()
    srv.cancel(1)
    while not srv.idle():
        srv.round()
    # a non-finite state fails at its first check
    assert srv.results[0].status == "failed" and srv.results[0].steps == 20
    assert srv.results[1].status == "cancelled" and srv.results[1].steps == 20
    assert srv.results[2].status == "step_cap" and srv.results[2].steps == 200
    assert srv.used == 0 and not srv.active
    srv.close()
    from cavityflow_cuda_amd.serving import solve_sequential
    alone = solve_sequential([SolveRequest(0, size, 0.1, 100.0, max_steps=200, initial=bad)],
                             check_every=20)
    assert alone[0].status == "failed" and alone[0].steps == 20


def test_open_loop_run_reports_latency(hip):
    from cavityflow_cuda_amd.serving import CavityServer, latency_summary, SolveRequest
    print("[case] serving open loop", flush=True)
    reqs = [SolveRequest(i, s, u, re, arrival=0.02 * i, max_steps=300)
            for i, (s, u, re) in enumerate(MIX)]
    srv = CavityServer("f32")
    res = srv.run(reqs)
    s = latency_summary(res)
    assert s["completed"] == len(MIX)
    assert 0 < s["p50_s"] <= s["p99_s"] <= s["max_s"]
    assert all(r.admitted >= r.arrival for r in res.values())
    srv.close()


def test_srpt_runs_least_remaining_work_first(hip):
    from cavityflow_cuda_amd.serving import CavityServer, SolveRequest, solve_sequential
    print("[case] serving srpt order", flush=True)
    reqs = [SolveRequest(0, (30, 30, 30), 0.1, 300.0, max_steps=400, keep_fields=True),
            SolveRequest(1, (12, 10, 8), 0.1, 100.0, max_steps=400, keep_fields=True),
            SolveRequest(2, (20, 16, 12), 0.1, 200.0, max_steps=200, keep_fields=True)]
    # one member per round: always the one with the least cells x steps left
    srv = CavityServer("f32", policy="srpt", min_batch_cells=1)
    for r in reqs:
        srv.submit(r)
    order = []
    while not srv.idle():
        for r in srv.round():
            order.append(r.id)
    assert order == [1, 2, 0]
    alone = solve_sequential([SolveRequest(r.id, r.size, r.u_lid, r.reynolds, max_steps=r.max_steps,
                                           keep_fields=True) for r in reqs])
    for i in range(3):
        assert srv.results[i].steps == alone[i].steps
        assert srv.results[i].fields.tobytes() == alone[i].fields.tobytes(), i
    srv.close()


def test_srpt_oldest_keeps_the_oldest_solve_moving(hip):
    from cavityflow_cuda_amd.serving import CavityServer, SolveRequest
    print("[case] serving srpt_oldest", flush=True)
    srv = CavityServer("f32", policy="srpt_oldest", min_batch_cells=1)
    srv.submit(SolveRequest(0, (30, 30, 30), 0.1, 300.0, max_steps=400))
    srv.round()                              # request 0 is the oldest resident solve
    for i in range(1, 4):
        srv.submit(SolveRequest(i, (10, 10, 10), 0.1, 100.0, max_steps=200))
    steps0 = []
    while not srv.idle():
        srv.round()
        a0 = [a for a in srv.active if a[0].id == 0]
        if a0:
            steps0.append(a0[0][3])
    # the big, oldest solve advanced every round while the small ones ran
    assert steps0 == sorted(steps0) and len(set(steps0)) == len(steps0)
    assert all(srv.results[i].status == "step_cap" for i in range(4))
    srv.close()
