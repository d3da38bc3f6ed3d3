"""Checkpoint of the whole A-A state (lbm_save_state / lbm_load_state): a run
resumed from a checkpoint is bit-identical to the uninterrupted run, for
every storage precision and at both step parities."""
import pytest

pytestmark = pytest.mark.gpu

SIZE = (18, 11, 14)


def make(hip, precision, lid=True):
    p = hip.cavity_parameters(SIZE, 0.1, 150.0, precision)
    s = hip.LbmSolverHIP(SIZE, precision=precision, inv_tau=p["inv_tau"],
                         lid_velocity=p["velocity_dimless"][0])
    if lid:
        s.setDrivenCavitySzenario()
    return s


def advance(s, first, count):
    for i in range(first, first + count):
        (s.simulationStepBeta if i & 1 else s.simulationStepAlpha)()


@pytest.mark.parametrize("precision", ["f32", "f64", "f16s", "bf16s", "f16c"])
@pytest.mark.parametrize("split", [40, 41])
def test_resume_equals_uninterrupted_run(hip, precision, split):
    print("[case] checkpoint", precision, split, flush=True)
    total = 90
    a = make(hip, precision)
    advance(a, 0, total)
    b = make(hip, precision)
    advance(b, 0, split)
    state = b.save_state()
    b.close()
    c = make(hip, precision, lid=False)       # flags come with the state
    c.load_state(state)
    advance(c, split, total - split)
    assert c.save_state().tobytes() == a.save_state().tobytes()
    assert c.getMoments(0).tobytes() == a.getMoments(0).tobytes()
    a.close()
    c.close()


def test_state_size_and_mismatch(hip):
    from cavityflow_cuda_amd import LbmError
    print("[case] checkpoint errors", flush=True)
    s = make(hip, "f16s")
    st = s.save_state()
    n = (SIZE[0] + 2) * (SIZE[1] + 2) * (SIZE[2] + 2)
    assert st.size == n * (19 * 2 + 1)
    assert (st[19 * 2 * n:] != 0).all()               # every cell has a flag
    with pytest.raises(LbmError):
        s.load_state(st[:-1])
    s.close()


def test_served_request_resumes_from_checkpoint(hip):
    from cavityflow_cuda_amd.serving import CavityServer, SolveRequest, solve_sequential
    print("[case] checkpoint serving", flush=True)
    b = make(hip, "f32")
    advance(b, 0, 100)
    state = b.save_state()
    b.close()
    srv = CavityServer("f32")
    srv.submit(SolveRequest(0, SIZE, 0.1, 150.0, max_steps=200, keep_fields=True, initial=state))
    srv.submit(SolveRequest(1, (9, 9, 9), 0.1, 80.0, max_steps=300, keep_fields=True))
    while not srv.idle():
        srv.round()
    straight = solve_sequential([SolveRequest(0, SIZE, 0.1, 150.0, max_steps=300,
                                              keep_fields=True)])
    assert srv.results[0].fields.tobytes() == straight[0].fields.tobytes()
    srv.close()
