"""Host logic of the solve server that needs no GPU: request draws, the
device-footprint query and admission-time rejections."""
import math

import pytest


def test_draw_requests_is_seeded_and_in_range(hip):
    from cavityflow_cuda_amd.serving import draw_requests
    a, ra = draw_requests(40, 7, rate=2.0, max_steps=500)
    b, rb = draw_requests(40, 7, rate=2.0, max_steps=500)
    assert ra == rb and [x.size for x in a] == [x.size for x in b]
    arr = [x.arrival for x in a]
    assert arr == sorted(arr) and arr[0] > 0
    for x in a:
        assert 48 <= min(x.size) and max(x.size) <= 384
        assert max(x.size) <= 2 * min(x.size)
        assert 100 <= x.reynolds <= 3200 and x.u_lid == 0.1
        assert 0.51 <= x.params["tau"] <= 1.9


def test_device_bytes_matches_allocation_formula(hip):
    from cavityflow_cuda_amd.solver import device_bytes
    n = 34 * 34 * 34
    stride = (n + 255) // 256 * 256
    partials = (1024 + 1) * 5 * 8 + 64           # block partials, results, overflow word
    assert device_bytes((32, 32, 32), "f32") == 19 * stride * 4 + stride + partials + 12 * stride
    assert device_bytes((32, 32, 32), "f16s") == 19 * stride * 2 + stride + partials + 12 * stride
    assert (device_bytes((32, 32, 32), "f32", pair_buffer=True)
            == device_bytes((32, 32, 32), "f32") + 19 * stride * 4)
    assert device_bytes((32, 32, 32), "f64", True) == (19 * stride * 8 + stride + partials
                                                      + 8 * stride + 24 * stride + 24 * stride)
    from cavityflow_cuda_amd import LbmError
    with pytest.raises(LbmError):
        device_bytes((0, 4, 4))


def test_submit_rejections(hip):
    from cavityflow_cuda_amd.serving import CavityServer, SolveRequest
    from cavityflow_cuda_amd.solver import device_bytes
    srv = CavityServer(hbm_budget=device_bytes((64, 64, 64)) + 1)
    srv.submit(SolveRequest(0, (64, 64, 64), 0.1, 500.0, max_steps=250))      # not a multiple of K
    srv.submit(SolveRequest(1, (32, 32, 32), 0.02, 2000.0 * 40))              # adaptation fails
    srv.submit(SolveRequest(2, (65, 64, 64), 0.1, 500.0))                      # over the budget
    srv.submit(SolveRequest(3, (64, 64, 64), 0.1, 500.0))                      # fits exactly
    assert [srv.results[i].status for i in range(3)] == ["rejected"] * 3
    assert "multiple" in srv.results[0].message and "budget" in srv.results[2].message
    assert len(srv.queue) == 1 and srv.queue[0].id == 3 and 3 not in srv.results
    # 1030^3 FP32 cells: a 4.37 GB plane, beyond the batch kernels' 32-bit offsets
    big = CavityServer(hbm_budget=1e12)
    big.submit(SolveRequest(4, (1028, 1028, 1028), 0.1, 3000.0))
    big.submit(SolveRequest(5, (1000, 1000, 1000), 0.1, 3000.0))
    assert big.results[4].status == "rejected"
    assert "too large" in big.results[4].message
    assert len(big.queue) == 1 and big.queue[0].id == 5


def test_latency_summary():
    from cavityflow_cuda_amd.serving import SolveResult, latency_summary
    res = {i: SolveResult(i, "step_cap", arrival=float(i), finished=float(i) + 1.0 + (i == 9))
           for i in range(10)}
    res[10] = SolveResult(10, "rejected", arrival=0.0, finished=0.0)
    s = latency_summary(res)
    assert s["completed"] == 10 and s["p50_s"] == 1.0 and s["max_s"] == 2.0
    assert math.isclose(s["span_s"], 11.0) and math.isclose(s["solves_per_s"], 10 / 11.0)


def test_hrrn_orders_by_response_ratio(hip):
    """policy="hrrn": (waited + remaining work at the measured rate) /
    remaining work, highest first; a long solve that has waited long enough
    goes ahead of a short one that just arrived."""
    from cavityflow_cuda_amd.serving import CavityServer, SolveRequest
    srv = CavityServer("f32", policy="hrrn", min_batch_cells=10**12)
    srv.clock = lambda: 10.0
    srv.cell_updates = 10**9                     # 1e8 cell updates per second so far
    big = SolveRequest(0, (100, 100, 100), 0.1, 100.0, arrival=0.0, max_steps=1000)
    fresh = SolveRequest(1, (10, 10, 100), 0.1, 100.0, arrival=9.95, max_steps=1000)
    waited = SolveRequest(2, (10, 10, 100), 0.1, 100.0, arrival=9.0, max_steps=1000)
    # [request, solver, bytes, steps done, admitted]; the solver is not used
    srv.active = [[big, None, 0, 0, 0.0], [fresh, None, 0, 0, 9.95], [waited, None, 0, 0, 9.0]]
    # ratios: waited (1 + 0.1) / 0.1 = 11, big (10 + 10) / 10 = 2, fresh (0.05 + 0.1) / 0.1 = 1.5
    assert [a[0].id for a in srv._select()] == [2, 0, 1]
    srv.min_batch_cells = 1
    assert [a[0].id for a in srv._select()] == [2]
    srv.active = []
    with pytest.raises(ValueError):
        CavityServer("f32", policy="fifo")
