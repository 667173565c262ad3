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
    """Every layer's weight and bias gradient (MFMA backward convolutions
    with BF16 operands, FP32 accumulation) against the oracle rounding the
    same tensors to BF16 at the same points, within 3e-2 relative per
    layer; the loss within 1e-3."""
    P = pcnet.init_params(4)
    net = learned.PCNet()
    net.set_params(learned.flatten(P))
    fields, target = _rand_case(seed=1)
    L, g = net.gradient(fields, target)
    # the GPU packs FP32 parameters, so the oracle starts from those
    P32 = {k: v.astype(np.float32).astype(np.float64) for k, v in P.items()}
    y_ref, C = pcnet.forward(P32, pcnet.make_input(fields.astype(np.float64)), rounded=True)
    L_ref, gy = pcnet.loss(y_ref, pcnet.pool4_target(target))
    G = pcnet.backward(P32, C, gy)
    assert L == pytest.approx(L_ref, rel=1e-3)
    Gg = learned.unflatten(g)
    worst = {}
    for name, _, _ in learned.LAYERS:
        for part in (".w", ".b"):
            a, b = Gg[name + part].ravel(), G[name + part].ravel()
            assert np.linalg.norm(b) > 0, name + part
            worst[name + part] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    print("per-layer relative gradient error:", {k: f"{v:.2e}" for k, v in worst.items()}, flush=True)
    for k, v in worst.items():
        assert v <= 3e-2, (k, v)
    net.close()


def test_first_stage_gate_keeps_the_answer_and_the_cost(gpu):
    """The learned first stage of the mixed solver (cosine loss on the
    solver's own corrections, two 64^3 realisations), on a held-out
    realisation: the FP64 answer is unchanged, and the gate, which runs the
    stage only while it saves solve time, keeps the iteration total within
    5% of the solve without it (the stage gains nothing measurable here:
    the multigrid V-cycle already removes the smooth error it predicts)."""
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
    assert its1 <= 1.05 * its0
    net.close()


def test_energy_training_gains_inside_flexible_cg(gpu):
    """Energy-norm training end to end on the GPU: stage-point pairs of
    flexible-CG solves on eight 64^3 realisations (the device training set,
    with each solve's operator), 15 epochs, then CG + Jacobi (the
    reference's method) against flexible CG with the trained stage on a
    held-out realisation: the same FP64 answer with at most 80% of the
    iterations over steps 4..12."""
    net = learned.PCNet(seed=1)
    npairs = learned.build_energy_set(net, nreal=8, grid=(64, 64, 64), nsteps=20, every=4, start=4,
                                      stride=2, seed0=500)
    losses = net.train_energy(epochs=15, lr=1e-3, seed=0)
    print(f"{npairs} pairs; energy loss per epoch:", np.round(losses, 4).tolist(), flush=True)
    assert losses[-1] < 0.7 * losses[0]
    case = io.workload("plain-small", seed=1001, grid=(64, 64, 64), corr_len=4.0)
    r = learned.compare(case, net, 12, skip=3, solver="cg", extrapolate=False, fcg=True)
    its0, its1 = sum(r["plain"]["iterations"]), sum(r["net"]["iterations"])
    print(f"steps 4..12: CG + Jacobi {r['plain']['iterations']} ({r['plain']['seconds']:.3f} s), "
          f"flexible CG with the stage {r['net']['iterations']} ({r['net']['seconds']:.3f} s, "
          f"{r['net']['learned']} stage applications)", flush=True)
    p0, p1 = r["plain"]["pressure"], r["net"]["pressure"]
    assert np.abs(p1 - p0).max() <= 1e-7 * np.abs(p0).max()
    assert r["net"]["learned"] > 0
    assert its1 <= 0.8 * its0
    net.close()


def test_learned_stage_inside_flexible_cg_gain(gpu):
    """The shipped stage weights (energy-norm training at 192^3 on stage
    points captured every 50 steps over 1000-step runs of 16 realisations)
    inside flexible CG with the Jacobi preconditioner, against CG + Jacobi
    (the reference's KSP settings) at 192^3 on a realisation they were not
    trained on, 12 steps from the previous pressure: the answers agree, and
    over steps 4..12 (the first three include 1000-iteration capped solves
    of CG + Jacobi) the learned path needs at most 70% of the iterations
    (62% measured) and at most 90% of the solve time (75% measured)."""
    net = learned.PCNet()
    net.load(learned.WEIGHTS_FCG)
    case = io.workload("plain-small", seed=1001, grid=(192, 192, 192), corr_len=4.0)
    r = learned.compare(case, net, 12, skip=3, solver="cg", extrapolate=False, fcg=True)
    its0, its1 = r["plain"]["iterations"], r["net"]["iterations"]
    t0, t1 = r["plain"]["seconds"], r["net"]["seconds"]
    print(f"iterations per step, CG + Jacobi: {its0} ({t0:.3f} s)", flush=True)
    print(f"iterations per step, flexible CG with the learned stage: {its1} ({t1:.3f} s, "
          f"{r['net']['learned']} stage applications)", flush=True)
    p0, p1 = r["plain"]["pressure"], r["net"]["pressure"]
    assert np.abs(p1 - p0).max() <= 1e-7 * np.abs(p0).max()
    assert sum(its1) <= 0.70 * sum(its0)
    assert t1 <= 0.90 * t0
    net.close()


def test_flexible_cg_with_the_stage_against_the_oracle(gpu):
    """The answer of flexible CG with the shipped learned stage, on the
    64x32x32 workload for 8 steps solving tightly, against the FP64 oracle
    (CG + Jacobi): every field within 1e-9 relative."""
    from oracle.oracle import OracleState
    net = learned.PCNet()
    net.load(learned.WEIGHTS_FCG)
    case = io.workload("plain-small", seed=1)
    g = WPFAW3D(case)
    g.attach_net(net)
    o = solver_opts("cg", rtol=1e-13, maxit=20000, flags=_lib.FLAG_LEARNED_FCG)
    stages = sum(g.step(o)["learned"] for _ in range(8))
    assert stages > 0
    ref = OracleState(case)
    ref.run(8, rtol=1e-13, maxit=20000)
    for name in ("pressure", "phase", "concent"):
        a, b = getattr(g, name)[1:-1, 1:-1, 1:-1], getattr(ref, name)[1:-1, 1:-1, 1:-1]
        err = float(np.abs(a - b).max() / np.abs(b).max())
        assert err <= 1e-9, (name, err)
    g.close()
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


def test_energy_gradient_matches_oracle(gpu):
    """The energy-norm loss (share of the error's A-norm the line search
    along the upsampled output leaves) and every layer's gradient, against
    the BF16-rounded oracle, on a real operator of the 64^3 workload."""
    from test_pcnet_oracle import _energy_case
    st, faces, fixed = _energy_case(n=64, seed=8)
    faces = faces / faces.max()
    P = pcnet.init_params(6)
    net = learned.PCNet()
    net.set_params(learned.flatten(P))
    fields, _ = _rand_case(seed=2)
    rng = np.random.default_rng(3)
    t = np.cumsum(np.cumsum(rng.normal(size=fixed.shape), axis=0), axis=1)   # a rough error field
    t[fixed] = 0.0
    L, g = net.gradient_energy(fields, t, faces, fixed)
    P32 = {k: v.astype(np.float32).astype(np.float64) for k, v in P.items()}
    y_ref, C = pcnet.forward(P32, pcnet.make_input(fields.astype(np.float64)), rounded=True)
    L_ref, gy = pcnet.energy_loss(y_ref, t.astype(np.float32).astype(np.float64),
                                  faces.astype(np.float32).astype(np.float64), fixed)
    G = pcnet.backward(P32, C, gy)
    assert 0.0 <= L_ref <= 1.0
    assert L == pytest.approx(L_ref, rel=1e-3, abs=1e-6)
    Gg = learned.unflatten(g)
    worst = {}
    for name, _, _ in learned.LAYERS:
        for part in (".w", ".b"):
            a, b = Gg[name + part].ravel(), G[name + part].ravel()
            worst[name + part] = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
    print(f"energy loss {L:.6f} (oracle {L_ref:.6f}); per-layer relative gradient error:",
          {k: f"{v:.2e}" for k, v in worst.items()}, flush=True)
    for k, v in worst.items():
        assert v <= 3e-2, (k, v)
    net.close()
