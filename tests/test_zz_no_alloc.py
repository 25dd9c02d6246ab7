"""The allocation-free sweep (moc_solver_transport_sweep_noalloc) against the
sweep it is a variant of (moc_solver_transport_sweep), through the C ABI.

Inputs and outputs are device buffers (torch tensors used as plain device
memory); the entry point's own buffers live in the library's static scratch
array, which every call on the GPU shares and which keeps the last call's
contents.  Calls in a row (the second on a smaller problem than the first) check that
nothing of an earlier call leaks into a later one.

Bars: the 7-group sweeps (LDS sweep, exact int64 tally) equal the host-buffer
sweep bit for bit (tally and outgoing angular fluxes); the 3D z-stack sweep
(exact int64 tally too) likewise; the 70-group sweep (FP32 atomics, whose
order varies from run to run; the host-buffer form sums per segment in FP64)
within 1e-6 relative L2 for the tally and bit for bit for the angular
fluxes."""
import ctypes as C

import numpy as np
import pytest

import openmoc_amd as om
from openmoc_amd._lib import check, lib

MOC_ERR_INVALID = -1

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def log(msg):
    print(msg, flush=True)


def _inputs(tg, seed):
    N, G, T = tg.getNumFSRs(), tg.getNumEnergyGroups(), tg.getNumTracks()
    P = 1 if tg.info.is_3d else tg.info.num_polar // 2
    rng = np.random.default_rng(seed)
    Q = rng.uniform(0.01, 1.0, N * G)
    psi = rng.uniform(0.0, 2.0, 2 * T * P * G).astype(np.float32)
    return Q, psi


def _noalloc(solver, Q, psi, stream=None):
    dev = torch.device("cuda", 0)
    dQ = torch.from_numpy(Q).to(dev)
    dpsi = torch.from_numpy(psi).to(dev)
    dout = torch.full_like(dpsi, -1.0)
    dtally = torch.full((Q.size,), -1.0, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    check(lib().moc_solver_transport_sweep_noalloc(
        solver._h, C.c_void_p(dQ.data_ptr()), C.c_void_p(dpsi.data_ptr()),
        C.c_void_p(dout.data_ptr()), C.c_void_p(dtally.data_ptr()), stream))
    torch.cuda.synchronize()
    return dtally.cpu().numpy(), dout.cpu().numpy()


def _scratch_bytes(solver):
    b = C.c_int64()
    check(lib().moc_solver_noalloc_scratch_bytes(solver._h, C.byref(b)))
    return b.value


def test_noalloc_sweep_equals_transport_sweep_on_calls_in_a_row(tracks):
    cases = [("lattice_17x17_7g", 11), ("lattice_3x3_vac_7g", 12), ("pin_cell_7g", 13),
             ("pin_cell_3d_7g", 14)]
    sizes = []
    for name, seed in cases:
        tg = tracks(name)
        s = om.HipSolver(tg)
        need = _scratch_bytes(s)
        sizes.append(need)
        Q, psi = _inputs(tg, seed)
        log(f"no-alloc sweep: {name}, {need} bytes of scratch, shape {s.getLaunchShape()}")
        t_ref, p_ref = s.transportSweep(Q, psi)
        t, p = _noalloc(s, Q, psi)
        np.testing.assert_array_equal(t, t_ref)
        np.testing.assert_array_equal(p, p_ref)
        # the same call again on the same inputs gives the same bits
        t2, p2 = _noalloc(s, Q, psi)
        np.testing.assert_array_equal(t2, t_ref)
        np.testing.assert_array_equal(p2, p_ref)
    assert sizes[1] < sizes[0]


def test_noalloc_sweep_70_groups_and_the_size_limit(tracks):
    tg = tracks("pin_cell_70g")
    s = om.HipSolver(tg)
    # the global-tally sweep: the allocation-free form adds FP32 atomics into
    # its scratch tally, the host-buffer form stores per segment
    s.setLaunchShape(lds_groups=0, groups_per_lane=2, lanes_per_track=35, tracks_per_wave=1)
    Q, psi = _inputs(tg, 21)
    t_ref, p_ref = s.transportSweep(Q, psi)
    t, p = _noalloc(s, Q, psi)
    rl = float(np.linalg.norm(t - t_ref) / np.linalg.norm(t_ref))
    log(f"no-alloc sweep, 70 groups: tally rel L2 {rl:.2e}")
    assert rl < 1e-6
    np.testing.assert_array_equal(p, p_ref)
    # a problem beyond the scratch array is refused before anything runs
    big = om.HipSolver(tracks("lattice_17x17_3d_7g"))
    assert _scratch_bytes(big) > (64 << 20)
    Qb, psib = _inputs(tracks("lattice_17x17_3d_7g"), 22)
    dev = torch.device("cuda", 0)
    dQ = torch.from_numpy(Qb).to(dev)
    dpsi = torch.from_numpy(psib).to(dev)
    dt = torch.empty(Qb.size, dtype=torch.float64, device=dev)
    rc = lib().moc_solver_transport_sweep_noalloc(
        big._h, C.c_void_p(dQ.data_ptr()), C.c_void_p(dpsi.data_ptr()),
        C.c_void_p(dpsi.data_ptr()), C.c_void_p(dt.data_ptr()), None)
    assert rc == MOC_ERR_INVALID
    assert b"scratch" in lib().moc_last_error()
