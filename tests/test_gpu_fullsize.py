"""Parity at the benchmark's full size (512^3 interior, FP32, Re 1000, lid
0.1) through size-independent properties: a 2x1x1 split with on-device face
exchange equals the monolithic run bit for bit (the reference's own
validation semantic, main.cpp:121-303), a checkpoint resumes bit-identically,
and the flow keeps the cavity's mirror symmetry in z."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 512
STEPS = 40


def say(*a):
    print("[case] fullsize", *a, flush=True)


@pytest.fixture(scope="module")
def mono(hip):
    say("monolithic", N, STEPS)
    c = hip.CavityController((N, N, N), u_lid=0.1, reynolds=1000.0, precision="f32",
                             store_outputs=True)
    for _ in range(STEPS):
        c.computeNextStep()
    c.sync()
    u = c.getVelocities()
    yield c, u
    c.close()


def test_split_equals_monolithic_at_full_size(hip, mono):
    say("split 2x1x1")
    _, u_mono = mono
    s = hip.CavityController((N, N, N), u_lid=0.1, reynolds=1000.0, precision="f32",
                             store_outputs=True, split=(2, 1, 1))
    for _ in range(STEPS):
        s.computeNextStep()
    s.sync()
    u = s.getVelocities()
    s.close()
    assert u.tobytes() == u_mono.tobytes()
    assert np.abs(u).max() > 0.05          # the lid has set the top layers moving


def test_mirror_symmetry_in_z(mono):
    say("mirror symmetry")
    _, u = mono
    ux, uy, uz = u[0], u[1], u[2]           # [z][y][x]
    scale = float(np.abs(ux).max())
    assert scale > 0.05
    for comp, sign in ((ux, 1.0), (uy, 1.0), (uz, -1.0)):
        d = np.abs(comp - sign * comp[::-1]).max()
        assert d <= 1e-5 * scale, d


def test_checkpoint_resume_at_full_size(hip, mono):
    say("checkpoint")
    c, _ = mono
    s = next(iter(c.solvers.values()))
    state = s.save_state()                  # ~10.4 GB of populations + flags
    p = hip.cavity_parameters((N, N, N), 0.1, 1000.0, "f32")
    r = hip.LbmSolverHIP((N, N, N), precision="f32", inv_tau=p["inv_tau"],
                         lid_velocity=p["velocity_dimless"][0], store_densities=False,
                         store_velocities=False)
    r.load_state(state)
    del state
    for n in range(4):
        (s.simulationStepBeta if n & 1 else s.simulationStepAlpha)()
        (r.simulationStepBeta if n & 1 else r.simulationStepAlpha)()
    a = s.getMoments(0)
    b = r.getMoments(0)
    r.close()
    assert a.tobytes() == b.tobytes()


def test_64bit_indexing_domain_equals_split(hip):
    """A 1024^3 cavity (1026^3 cells: each population plane 4.32 GB, beyond
    32-bit byte offsets, so the step kernels take their 64-bit indexing path)
    against the same cavity split 2x1x1 (planes of 2.16 GB: the 32-bit path),
    bit for bit after 8 steps with the on-device face exchange; and against
    itself with its x faces exchanged inside the step kernels."""
    n, steps = 1024, 8
    say("64-bit indexing", n, steps)
    c = hip.CavityController((n, n, n), u_lid=0.1, reynolds=1000.0, precision="f32",
                             store_outputs=True)
    arr = next(iter(c.solvers.values())).device_arrays()
    assert arr.plane_stride * arr.element_bytes >= 1 << 32
    for _ in range(steps):
        c.computeNextStep()
    c.sync()
    u_mono = c.getVelocities()
    c.close()
    s = hip.CavityController((n, n, n), u_lid=0.1, reynolds=1000.0, precision="f32",
                             store_outputs=True, split=(2, 1, 1))
    for sub in s.solvers.values():
        a = sub.device_arrays()
        assert a.plane_stride * a.element_bytes < 1 << 32
    for _ in range(steps):
        s.computeNextStep()
    s.sync()
    u = s.getVelocities()
    s.close()
    assert np.abs(u_mono).max() > 0.05
    assert u.tobytes() == u_mono.tobytes()
    # the same cavity receiving its own x faces through the fused pack and
    # unpack (two buffer sets swapping every alpha + beta pair): with 64-bit
    # indexing FP32 beta then defaults to 512-thread workgroups, the
    # 1024-thread kernel would spill
    import ctypes
    from hiprt import DeviceBuffer
    from cavityflow_cuda_amd._lib import (LAUNCH_NONTEMPORAL, LAUNCH_WAVES3, LAUNCH_WIDE, check,
                                          lib)
    say("64-bit indexing, self-exchanged x faces", n, steps)
    f = hip.CavityController((n, n, n), u_lid=0.1, reynolds=1000.0, precision="f32",
                             store_outputs=True)
    sv = next(iter(f.solvers.values()))
    ne = sv.face_elements(0)
    bufs = [DeviceBuffer(ne * 4) for _ in range(4)]
    sets = [bufs[:2], bufs[2:]]
    for side in (0, 1):
        sv.pack_face(side, 0, sets[1][side])
        sv.pack_face(side, 1, sets[1][side])
    v = ctypes.c_int()
    for step in range(steps):
        j = (step >> 1) & 1
        sv.set_step_pack(*sets[j])
        sv.set_step_unpack(*sets[1 - j])
        if step == 0:
            for k, want in ((0, LAUNCH_NONTEMPORAL | LAUNCH_WAVES3), (1, LAUNCH_WIDE | LAUNCH_WAVES3)):
                check(lib().lbm_get_launch_options(sv.handle, k, ctypes.byref(v)))
                assert v.value == want, k
        f.computeNextStep()
    f.sync()
    u_fused = f.getVelocities()
    sv.set_step_pack(None, None)
    sv.set_step_unpack(None, None)
    f.close()
    for b in bufs:
        b.free()
    assert u_fused.tobytes() == u_mono.tobytes()
