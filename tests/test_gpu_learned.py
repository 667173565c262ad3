"""The learned pressure correction on the GPU: forward and gradients of the
hand-written BF16/MFMA network against the FP64 numpy oracle
(oracle/pcnet.py), training on solver-generated data, and what the trained
network gains inside the solve over the same solve without it."""
import ctypes

import numpy as np
import pytest

from mizuho_edit_amd import _lib, io, learned
from mizuho_edit_amd.wpfaw3d import WPFAW3D, solver_opts
from oracle import pcnet

pytestmark = pytest.mark.gpu


def _rand_case(n=64, seed=0):
    rng = np.random.default_rng(seed)
    fields = rng.normal(size=(3, n, n, n)).astype(np.float32)
    target = rng.normal(size=(n, n, n)).astype(np.float32)
    # smooth the target a little so the pooled target is not pure noise
    return fields, target


def test_forward_matches_oracle(gpu):
    P = pcnet.init_params(3)
    rng = np.random.default_rng(5)
    for k in P:
        if k.endswith(".b"):
            P[k] = rng.normal(0, 0.05, P[k].shape)
    net = learned.PCNet()
    net.set_params(learned.flatten(P))
    fields, _ = _rand_case()
    y = net.forward(fields)
    y_ref, _ = pcnet.forward(P, pcnet.make_input(fields))
    y_ref = y_ref[..., 0]
    # BF16 weights and activations through eight layers: ~1e-2 relative
    err = np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref)
    assert err < 3e-2, err
    net.close()


def test_gradient_matches_oracle(gpu):
    P = pcnet.init_params(4)
    net = learned.PCNet()
    net.set_params(learned.flatten(P))
    fields, target = _rand_case(seed=1)
    L, g = net.gradient(fields, target)
    y_ref, C = pcnet.forward(P, pcnet.make_input(fields))
    L_ref, gy = pcnet.loss(y_ref, pcnet.pool4_target(target))
    G = pcnet.backward(P, C, gy)
    assert L == pytest.approx(L_ref, rel=2e-2)
    Gg = learned.unflatten(g)
    for name, _, _ in learned.LAYERS:
        for part in (".w", ".b"):
            a, b = Gg[name + part].ravel(), G[name + part].ravel()
            if np.linalg.norm(b) < 1e-12:
                continue
            cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b))
            rel = np.linalg.norm(a - b) / np.linalg.norm(b)
            assert cos > 0.99 and rel < 0.15, (name + part, cos, rel)
    net.close()


def test_training_and_learned_stage_gain(gpu):
    """Train on corrections the solver itself produced on two 64^3
    realisations, then solve a held-out realisation with and without the
    network: same FP64 answer, and the record of what the network saves."""
    samples = []
    for i, corr in enumerate((3.0, 6.0)):
        case = io.workload("plain-small", seed=200 + i, grid=(64, 64, 64), corr_len=corr)
        samples += learned.capture_samples(case, nsteps=24, every=2)
    net = learned.PCNet(seed=1)
    losses = net.train(samples, epochs=30, lr=1e-3)
    print("training loss per epoch:", np.round(losses, 4).tolist(), flush=True)
    assert losses[-1] < 0.7 * losses[0]

    case = io.workload("plain-small", seed=300, grid=(64, 64, 64), corr_len=4.0)
    flags = _lib.FLAG_EXTRAPOLATE
    runs = {}
    for use_net in (False, True):
        g = WPFAW3D(case)
        if use_net:
            _lib.lib().wpf_attach_net(g._ctx, net.handle)
        its, acc = [], 0
        for s in range(20):
            o = solver_opts("mixed", rtol=1e-10,
                            flags=flags | (_lib.FLAG_LEARNED_WARM if use_net else 0))
            rec = g.step(o)
            assert rec["converged"] == 1
            if s >= 3:
                its.append(rec["iterations"])
                acc += rec["learned"]
        runs[use_net] = (sum(its), acc, g.pressure)
        g.close()
    its0, _, p0 = runs[False]
    its1, acc1, p1 = runs[True]
    print(f"FP32 inner iterations over steps 4..20: without network {its0}, with network {its1}, "
          f"learned stage accepted {acc1}/17", flush=True)
    assert np.abs(p1 - p0).max() <= 1e-7 * np.abs(p0).max()
    assert acc1 >= 1
    net.close()


def test_learned_stage_inside_flexible_cg_gain(gpu):
    """The shipped weights as a stage inside flexible CG with the Jacobi
    preconditioner, against CG + Jacobi (the reference's KSP settings) on a
    realisation the weights were not trained on: every solve meets the same
    FP64 test, the answers agree, and from step 6 on (residuals like the
    training data) the solves need fewer iterations.  The first steps, where
    the stage can stall the iteration and plain CG takes over, are reported."""
    net = learned.PCNet()
    net.load()
    case = io.workload("plain-small", seed=300, grid=(64, 64, 64), corr_len=4.0)
    runs = {}
    for fcg in (False, True):
        g = WPFAW3D(case)
        if fcg:
            g.attach_net(net)
        o = solver_opts("cg", rtol=1e-10, flags=_lib.FLAG_LEARNED_FCG if fcg else 0)
        recs = [g.step(o) for _ in range(12)]
        assert all(r["converged"] == 1 for r in recs)
        runs[fcg] = (recs, g.pressure)
        g.close()
    its0 = [r["iterations"] for r in runs[False][0]]
    its1 = [r["iterations"] for r in runs[True][0]]
    stages = sum(r["learned"] for r in runs[True][0][5:])
    t0 = sum(r["seconds"] for r in runs[False][0][5:])
    t1 = sum(r["seconds"] for r in runs[True][0][5:])
    print(f"iterations per step, CG + Jacobi: {its0}", flush=True)
    print(f"iterations per step, flexible CG with the learned stage: {its1}", flush=True)
    print(f"steps 6..12: {sum(its0[5:])} vs {sum(its1[5:])} iterations ({stages} stage applications), "
          f"solve time {t0:.3f} s vs {t1:.3f} s", flush=True)
    p0, p1 = runs[False][1], runs[True][1]
    assert np.abs(p1 - p0).max() <= 1e-7 * np.abs(p0).max()
    assert stages >= 7 * 10
    assert sum(its1[5:]) < 0.97 * sum(its0[5:])
    net.close()


def test_flexible_cg_stage_point_capture(gpu):
    """The training pairs of the stage inside flexible CG: at most 16 per
    solve, the residual channel normalised to unit rms, and the first pair
    (the solve's start) equal to the start-of-solve capture."""
    net = learned.PCNet()
    net.load()
    case = io.workload("plain-small", seed=301, grid=(64, 64, 64), corr_len=4.0)
    g = WPFAW3D(case)
    g.attach_net(net)
    g.run(3, solver_opts("cg", rtol=1e-10))
    rec = g.step(solver_opts("cg", rtol=1e-10, flags=_lib.FLAG_LEARNED_FCG | _lib.FLAG_CAPTURE))
    assert rec["converged"] == 1 and rec["learned"] >= 2
    L = _lib.lib()
    cnt = ctypes.c_int32()
    assert L.wpf_capture_fcg_count(g._ctx, ctypes.byref(cnt)) == 0
    assert 1 <= cnt.value <= 16 and cnt.value == min(16, (rec["learned"] + 1) // 2)
    n = case.params.ncell
    f0, t0 = np.zeros(3 * n, np.float32), np.zeros(n, np.float32)
    fs, ts = np.zeros(3 * n, np.float32), np.zeros(n, np.float32)
    assert L.wpf_capture_fcg_get(g._ctx, 0, f0.ctypes.data_as(_lib.F), t0.ctypes.data_as(_lib.F)) == 0
    assert L.wpf_capture_get(g._ctx, fs.ctypes.data_as(_lib.F), ts.ctypes.data_as(_lib.F)) == 0
    assert np.array_equal(t0, ts)
    assert np.allclose(f0, fs, rtol=1e-6, atol=1e-6)
    for j in range(cnt.value):
        f, t = np.zeros(3 * n, np.float32), np.zeros(n, np.float32)
        assert L.wpf_capture_fcg_get(g._ctx, j, f.ctypes.data_as(_lib.F), t.ctypes.data_as(_lib.F)) == 0
        assert abs(np.sqrt(np.mean(f[2 * n:].astype(np.float64) ** 2)) - 1.0) < 1e-3
        assert np.isfinite(t).all() and t.any()
    assert L.wpf_capture_fcg_get(g._ctx, cnt.value, f0.ctypes.data_as(_lib.F),
                                 t0.ctypes.data_as(_lib.F)) == _lib.WPF_ERR_ARG
    g.close()
    net.close()
