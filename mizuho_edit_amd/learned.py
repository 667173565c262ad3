"""The learned pressure correction: host side of the network in
k_learned.hip (a 3-D conv encoder-decoder, BF16 on MFMA), its training data
and its training loop.

Training data are (input, correction) pairs the plain solver produces
itself: with WPF_FLAG_CAPTURE a solve keeps its start state as the network
input (log k, pressure, Jacobi-scaled residual) and the correction it went on
to find.  For the stage inside flexible CG the pairs come from the stage
points of CG + Jacobi solves (the iterate there and its distance to the final
solution).  Training runs on the GPU through the library's own forward,
backward and Adam kernels.
"""
from __future__ import annotations

import ctypes
import os
from typing import List, Optional, Tuple

import numpy as np

from . import _lib, io
from ._lib import check

LAYERS = [("e1", 8, 16), ("e2", 16, 32), ("e3", 32, 48), ("e4", 48, 64),
          ("d3", 112, 48), ("d2", 80, 32), ("d1", 48, 16), ("out", 16, 1)]
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "pcnet.npz")
# weights for the stage inside flexible CG (trained on its stage points)
WEIGHTS_FCG = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "pcnet_fcg.npz")


def _fp(a):
    return a.ctypes.data_as(_lib.F)


def flatten(P: dict) -> np.ndarray:
    """{name.w: (co, 27, ci), name.b: (co,)} -> the library's flat order."""
    parts = []
    for name, ci, co in LAYERS:
        parts.append(np.asarray(P[name + ".w"], dtype=np.float32).reshape(-1))
        parts.append(np.asarray(P[name + ".b"], dtype=np.float32).reshape(-1))
    return np.concatenate(parts)


def unflatten(v: np.ndarray) -> dict:
    P, off = {}, 0
    for name, ci, co in LAYERS:
        nw = co * 27 * ci
        P[name + ".w"] = v[off:off + nw].reshape(co, 27, ci).astype(np.float64)
        off += nw
        P[name + ".b"] = v[off:off + co].astype(np.float64)
        off += co
    return P


class PCNet:
    def __init__(self, seed: int = 0):
        h = ctypes.c_void_p()
        check(_lib.lib().wpf_net_create(seed, ctypes.byref(h)), "wpf_net_create")
        self._h = h
        self.nparams = int(_lib.lib().wpf_net_num_params())

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().wpf_net_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def get_params(self) -> np.ndarray:
        v = np.zeros(self.nparams, np.float32)
        check(_lib.lib().wpf_net_get_params(self._h, _fp(v), self.nparams), "get_params")
        return v

    def set_params(self, v: np.ndarray):
        v = np.ascontiguousarray(v, dtype=np.float32)
        check(_lib.lib().wpf_net_set_params(self._h, _fp(v), self.nparams), "set_params")

    def forward(self, fields: np.ndarray) -> np.ndarray:
        """fields: (3, nx, ny, nz) -> correction at (nx/4, ny/4, nz/4)."""
        _, nx, ny, nz = fields.shape
        f = np.ascontiguousarray(np.asarray(fields, np.float32).reshape(3, -1, order="F"))
        out = np.zeros((nx // 4) * (ny // 4) * (nz // 4), np.float32)
        check(_lib.lib().wpf_net_forward(self._h, nx, ny, nz, _fp(f), _fp(out)), "net_forward")
        return out.reshape((nx // 4, ny // 4, nz // 4), order="F")

    def gradient(self, fields, target) -> Tuple[float, np.ndarray]:
        _, nx, ny, nz = fields.shape
        f = np.ascontiguousarray(np.asarray(fields, np.float32).reshape(3, -1, order="F"))
        t = np.ascontiguousarray(np.asarray(target, np.float32).reshape(-1, order="F"))
        g = np.zeros(self.nparams, np.float32)
        L = ctypes.c_double()
        check(_lib.lib().wpf_net_gradient(self._h, nx, ny, nz, _fp(f), _fp(t), ctypes.byref(L), _fp(g)),
              "net_gradient")
        return L.value, g

    def train(self, samples: List[Tuple[np.ndarray, np.ndarray]], epochs: int = 10,
              lr: float = 1e-3) -> np.ndarray:
        """samples: [(fields (3,nx,ny,nz), target (nx,ny,nz))], one grid size."""
        _, nx, ny, nz = samples[0][0].shape
        F = np.ascontiguousarray(np.stack([np.asarray(f, np.float32).reshape(3, -1, order="F")
                                           for f, _ in samples]))
        T = np.ascontiguousarray(np.stack([np.asarray(t, np.float32).reshape(-1, order="F")
                                           for _, t in samples]))
        losses = np.zeros(epochs)
        check(_lib.lib().wpf_net_train(self._h, nx, ny, nz, len(samples), _fp(F), _fp(T), epochs,
                                       ctypes.c_float(lr), losses.ctypes.data_as(_lib.D)), "net_train")
        return losses

    def gradient_energy(self, fields, target, faces, fixed) -> Tuple[float, np.ndarray]:
        """Energy-norm loss and gradient of one pair: fields (3,nx,ny,nz),
        target (nx,ny,nz), faces (3,nx,ny,nz) (+x/+y/+z face of each cell, 0
        on the global edge), fixed (nx,ny,nz) bool (Dirichlet cells)."""
        _, nx, ny, nz = fields.shape
        f = np.ascontiguousarray(np.asarray(fields, np.float32).reshape(3, -1, order="F"))
        t = np.ascontiguousarray(np.asarray(target, np.float32).reshape(-1, order="F"))
        a = np.ascontiguousarray(np.asarray(faces, np.float32).reshape(3, -1, order="F"))
        m = np.ascontiguousarray(np.asarray(fixed, np.uint8).reshape(-1, order="F"))
        g = np.zeros(self.nparams, np.float32)
        L = ctypes.c_double()
        check(_lib.lib().wpf_net_gradient_energy(self._h, nx, ny, nz, _fp(f), _fp(t), _fp(a),
                                                 m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                                 ctypes.byref(L), _fp(g)), "net_gradient_energy")
        return L.value, g

    def dataset_reserve(self, grid, max_pairs: int, max_ops: int):
        check(_lib.lib().wpf_net_dataset_reserve(self._h, *grid, max_pairs, max_ops), "dataset_reserve")

    def dataset_add(self, g, stride: int = 1) -> int:
        """Every stride-th pair captured by g's last flexible-CG solve, with
        the operator of that solve, into the device training set."""
        n = ctypes.c_int32()
        check(_lib.lib().wpf_net_dataset_add(self._h, g._ctx, stride, ctypes.byref(n)), "dataset_add")
        return n.value

    def train_energy(self, epochs: int = 10, lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        losses = np.zeros(epochs)
        check(_lib.lib().wpf_net_train_energy(self._h, epochs, ctypes.c_float(lr), seed,
                                              losses.ctypes.data_as(_lib.D)), "net_train_energy")
        return losses

    def save(self, path: str = WEIGHTS):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.savez_compressed(path, params=self.get_params())

    def load(self, path: str = WEIGHTS):
        with np.load(path, allow_pickle=False) as z:
            self.set_params(z["params"])


def capture_samples(case, nsteps: int, every: int, opts_kw=None, start: int = 3):
    """Run the plain path on `case`, capturing (input, correction) every
    `every` steps from step `start` on."""
    from .wpfaw3d import WPFAW3D, solver_opts
    kw = dict(rtol=1e-10)
    kw.update(opts_kw or {})
    base = kw.pop("flags", _lib.FLAG_EXTRAPOLATE)
    g = WPFAW3D(case, opts=solver_opts(kw.pop("solver", "mixed"), flags=base, **kw))
    p = case.params
    n = p.ncell
    out = []
    for s in range(1, nsteps + 1):
        cap = s >= start and (s - start) % every == 0
        o = solver_opts("mixed", rtol=1e-10, flags=base | (_lib.FLAG_CAPTURE if cap else 0))
        g.step(o)
        if cap:
            f = np.zeros(3 * n, np.float32)
            t = np.zeros(n, np.float32)
            check(_lib.lib().wpf_capture_get(g._ctx, _fp(f), _fp(t)), "capture_get")
            if not t.any():
                continue  # the solve started converged: nothing to learn
            out.append((f.reshape((3, p.nx, p.ny, p.nz), order="F"),
                        t.reshape((p.nx, p.ny, p.nz), order="F")))
    g.close()
    return out


def _upsample_1d(nf: int):
    """The network's trilinear N/4 -> N upsampling along one axis (the
    cell-centred weights of k_upsample_out), as a sparse matrix."""
    import scipy.sparse as sp
    nc = nf // 4
    cc = np.clip((np.arange(nf) + 0.5) / 4 - 0.5, 0, nc - 1)
    i0 = np.minimum(cc.astype(int), nc - 1)
    i1 = np.minimum(i0 + 1, nc - 1)
    f = cc - i0
    rows = np.concatenate([np.arange(nf), np.arange(nf)])
    return sp.csr_matrix((np.concatenate([1 - f, f]), (rows, np.concatenate([i0, i1]))), shape=(nf, nc))


class CoarseSolve:
    """The best correction the network's output space holds: for a residual
    r, y = P (P^T A P)^-1 P^T r, the A-orthogonal projection of the error
    onto the trilinear upsamplings of N/4 fields.  A is the SPD operator of
    the solve (free cells, from the device's face coefficients), P the
    network's upsampling with Dirichlet rows zero."""

    def __init__(self, case, faces):
        import scipy.sparse as sp
        import scipy.sparse.linalg as spla
        p = case.params
        nx, ny, nz = p.nx, p.ny, p.nz
        n = nx * ny * nz
        lw = np.asarray(case.lw)[1:-1, 1:-1, 1:-1].reshape(-1, order="F")
        press = np.array([0.0] + [w.pressure for w in case.wells])
        fixed = (lw > 0) & (press[np.maximum(lw, 0)] != 0.0)
        t = [np.asarray(faces[d])[1:-1, 1:-1, 1:-1].reshape(-1, order="F") for d in range(3)]
        idx = np.arange(n)
        rows, cols, vals = [], [], []
        diag = np.zeros(n)
        for d, step in enumerate((1, nx, nx * ny)):
            a, b = idx[:-step] if step < n else idx[:0], idx[step:]
            tt = t[d][a]
            keep = tt != 0.0
            a, b, tt = a[keep], b[keep], tt[keep]
            diag[a] += tt
            diag[b] += tt
            free = ~fixed[a] & ~fixed[b]
            rows += [a[free], b[free]]
            cols += [b[free], a[free]]
            vals += [-tt[free], -tt[free]]
        diag[fixed] = 1.0
        A = sp.csr_matrix((np.concatenate(vals + [diag]), (np.concatenate(rows + [idx]), np.concatenate(cols + [idx]))),
                          shape=(n, n))
        P = sp.kron(sp.kron(_upsample_1d(nz), _upsample_1d(ny)), _upsample_1d(nx)).tocsr()
        self.P = sp.diags((~fixed).astype(float)) @ P
        self.D = diag
        self.lu = spla.splu((self.P.T @ A @ self.P).tocsc())
        self.shape = (nx, ny, nz)

    def target(self, fields):
        """The coarse-solve correction (full resolution) for the residual of
        a network input (its channel 2 is D^-1 r up to a scale)."""
        r = np.asarray(fields[2], np.float64).reshape(-1, order="F") * self.D
        y = self.P @ self.lu.solve(self.P.T @ r)
        return y.reshape(self.shape, order="F").astype(np.float32)


def capture_fcg_samples(case, net: "PCNet", nsteps: int, start: int = 3, coarse: bool = False):
    """Run `case` with CG + Jacobi and the learned stage inside flexible CG
    (`net`), keeping from step `start` on the pairs of every solve's stage
    points: the network input there and the correction from that iterate to
    the solve's final solution (coarse=True: the CoarseSolve correction of
    that residual instead)."""
    from .wpfaw3d import WPFAW3D, solver_opts
    g = WPFAW3D(case)
    g.attach_net(net)
    p = case.params
    n = p.ncell
    out = []
    for s in range(1, nsteps + 1):
        cap = s >= start
        faces = None
        if cap and coarse:
            # the operator of this step's solve: the permeability pass of
            # the current fields, which the step repeats
            g.calcPermeability()
            faces = g.download("face_t")
        g.step(solver_opts("cg", rtol=1e-10,
                           flags=_lib.FLAG_LEARNED_FCG | (_lib.FLAG_CAPTURE if cap else 0)))
        if not cap:
            continue
        cnt = ctypes.c_int32()
        check(_lib.lib().wpf_capture_fcg_count(g._ctx, ctypes.byref(cnt)), "capture_fcg_count")
        cs = CoarseSolve(case, faces) if coarse and cnt.value else None
        for j in range(cnt.value):
            f = np.zeros(3 * n, np.float32)
            t = np.zeros(n, np.float32)
            check(_lib.lib().wpf_capture_fcg_get(g._ctx, j, _fp(f), _fp(t)), "capture_fcg_get")
            if t.any():
                f = f.reshape((3, p.nx, p.ny, p.nz), order="F")
                out.append((f, cs.target(f) if cs else t.reshape((p.nx, p.ny, p.nz), order="F")))
    g.close()
    return out


def make_fcg_training_set(net: "PCNet", nreal: int = 8, grid=(64, 64, 64), nsteps: int = 12,
                          seed0: int = 100, coarse: bool = False):
    """Stage-point pairs from seeded realisations (correlation lengths
    2..8 cells), generated with `net` as the stage."""
    samples = []
    for i in range(nreal):
        corr = 2.0 + 6.0 * (i / max(1, nreal - 1))
        case = io.workload("plain-small", seed=seed0 + i, grid=grid, corr_len=corr)
        samples.extend(capture_fcg_samples(case, net, nsteps, coarse=coarse))
        print(f"  realisation {i + 1}/{nreal}: {len(samples)} pairs", flush=True)
    return samples


def make_training_set(nreal: int = 8, grid=(64, 64, 64), nsteps: int = 40, every: int = 4,
                      seed0: int = 100):
    """Seeded realisations of the plain workload's physics with correlation
    lengths spread over 2..8 cells."""
    samples = []
    for i in range(nreal):
        corr = 2.0 + 6.0 * (i / max(1, nreal - 1))
        case = io.workload("plain-small", seed=seed0 + i, grid=grid, corr_len=corr)
        samples.extend(capture_samples(case, nsteps, every))
    return samples


def energy_case(i: int, seed0: int, grid):
    """Realisation i of the training set: the plain workload's physics with
    a correlation length spread over 2..8 cells."""
    u = (i * 0.6180339887498949) % 1.0
    return io.workload("plain-small", seed=seed0 + i, grid=grid, corr_len=2.0 + 6.0 * u)


def build_energy_set(net: "PCNet", nreal: int, grid, nsteps: int, every: int, start: int = 4,
                     stride: int = 2, seed0: int = 500, log=print, pool=None, period: int = 16) -> int:
    """The energy-norm training set, kept on the device: seeded realisations
    of the plain workload's physics (correlation lengths 2..8 cells) are run
    with the fast mixed solver; at steps start, start + every, ... the step's
    solve is instead CG + Jacobi with `net` as the stage inside flexible CG,
    capturing its stage-point pairs, of which every stride-th goes into the
    set with that solve's operator.  Returns the pair count.  `pool` (a
    process pool started before the GPU was touched) builds the rocks ahead."""
    import functools
    from .wpfaw3d import WPFAW3D, solver_opts
    snaps = max(0, (nsteps - start) // every + 1)
    net.dataset_reserve(grid, max_pairs=nreal * snaps * ((16 + stride - 1) // stride), max_ops=nreal * snaps)
    fast = solver_opts("mixed", rtol=1e-10, inner_rtol=1e-4, mg_levels=6,
                       flags=_lib.FLAG_EXTRAPOLATE | _lib.FLAG_CHEBYSHEV | _lib.FLAG_OP16 | _lib.FLAG_PROJECT)
    cap = solver_opts("cg", rtol=1e-10, flags=_lib.FLAG_LEARNED_FCG | _lib.FLAG_CAPTURE)
    count = 0
    make = functools.partial(energy_case, seed0=seed0, grid=tuple(grid))
    cases = pool.imap(make, range(nreal)) if pool is not None else map(make, range(nreal))
    for i, case in enumerate(cases):
        g = WPFAW3D(case)
        g.attach_net(net)
        check(_lib.lib().wpf_set_fcg_period(g._ctx, period), "set_fcg_period")
        for s in range(1, nsteps + 1):
            if s >= start and (s - start) % every == 0:
                g.step(cap)
                count = net.dataset_add(g, stride)
            else:
                g.step(fast)
        g.close()
        log(f"  realisation {i + 1}/{nreal}: {count} pairs")
    return count


def compare(case, net: Optional["PCNet"], nsteps: int, skip: int = 3, solver: str = "mixed",
            inner_rtol: float = 1e-4, extrapolate: bool = True, fcg: bool = False, period: int = 16):
    """Run `case` for nsteps from the extrapolated start (or, with
    extrapolate=False, from the previous pressure), without and with the
    learned stage: the first stage, or with fcg=True (and solver "cg") the
    stage inside flexible CG.  Returns per-run iterations (steps after
    `skip`), learned stages (accepted first stages, or applications inside
    flexible CG), wall seconds and the final pressures."""
    import time
    from .wpfaw3d import WPFAW3D, solver_opts
    out = {}
    for use in ((False, True) if net is not None else (False,)):
        g = WPFAW3D(case)
        if use:
            g.attach_net(net)
            check(_lib.lib().wpf_set_fcg_period(g._ctx, period), "set_fcg_period")
        stage = _lib.FLAG_LEARNED_FCG if fcg else _lib.FLAG_LEARNED_WARM
        flags = (_lib.FLAG_EXTRAPOLATE if extrapolate else 0) | (stage if use else 0)
        o = solver_opts(solver, rtol=1e-10, inner_rtol=inner_rtol, flags=flags)
        its, acc, wall = [], 0, 0.0
        for s in range(nsteps):
            t0 = time.perf_counter()
            rec = g.step(o)
            dt = time.perf_counter() - t0
            if s >= skip:
                its.append(rec["iterations"])
                acc += rec["learned"]
                wall += dt
        out["net" if use else "plain"] = {"iterations": its, "learned": acc, "seconds": wall,
                                          "pressure": g.pressure}
        g.close()
    return out


def main(argv=None):
    """python -m mizuho_edit_amd.learned train|eval ..."""
    import argparse
    import json
    ap = argparse.ArgumentParser(prog="python -m mizuho_edit_amd.learned")
    ap.add_argument("mode", choices=("train", "train-fcg", "train-energy", "eval"))
    ap.add_argument("--weights", default=WEIGHTS)
    ap.add_argument("--nreal", type=int, default=16)
    ap.add_argument("--grid", type=int, default=64)
    ap.add_argument("--nsteps", type=int, default=40)
    ap.add_argument("--every", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-extrapolate", action="store_true",
                    help="eval: start each solve from the previous pressure, not the extrapolated one")
    ap.add_argument("--start", type=int, default=4, help="train-energy: first captured step")
    ap.add_argument("--period", type=int, default=16,
                    help="the learned stage every period-th preconditioner application of flexible CG")
    ap.add_argument("--stride", type=int, default=2,
                    help="train-energy: keep every stride-th captured stage point")
    ap.add_argument("--rounds", type=int, default=1,
                    help="train-energy: capture/train rounds (each round captures with the "
                         "network of the previous one)")
    ap.add_argument("--init", default=None, help="train-energy: start from these weights")
    ap.add_argument("--coarse", action="store_true",
                    help="train-fcg: the coarse-solve correction of each residual as the target")
    ap.add_argument("--fcg", action="store_true",
                    help="eval: the stage inside flexible CG with the Jacobi preconditioner "
                         "(against plain CG + Jacobi) instead of the first stage of the mixed solver")
    a = ap.parse_args(argv)
    n = a.grid
    if a.mode == "train-energy":
        import multiprocessing as mp
        # the rock generator runs in worker processes forked before this
        # process touches the GPU
        pool = mp.get_context("fork").Pool(min(8, os.cpu_count() or 1))
        net = PCNet(seed=a.seed)
        if a.init:
            net.load(a.init)
        for rnd in range(a.rounds):
            npairs = build_energy_set(net, a.nreal, (n, n, n), a.nsteps, a.every, start=a.start, stride=a.stride,
                                      seed0=500 + 1000 * rnd, pool=pool, period=a.period)
            print(f"round {rnd + 1}: {npairs} stage-point pairs at {n}^3", flush=True)
            losses = []
            for e in range(a.epochs):
                losses += net.train_energy(epochs=1, lr=a.lr, seed=e + 100 * rnd).tolist()
                print(f"  epoch {e + 1}/{a.epochs}: energy loss {losses[-1]:.4f}", flush=True)
            print("energy loss per epoch:", np.round(losses, 4).tolist(), flush=True)
        pool.close()
        out = a.weights if a.weights != WEIGHTS else WEIGHTS_FCG
        net.save(out)
        print(f"saved {out}", flush=True)
    elif a.mode == "train-fcg":
        # stage-point pairs generated with the first-stage weights as the
        # stage, then a network trained on them from scratch
        gen = PCNet()
        gen.load(WEIGHTS)
        samples = make_fcg_training_set(gen, a.nreal, (n, n, n), a.nsteps, coarse=a.coarse)
        gen.close()
        print(f"{len(samples)} stage-point samples at {n}^3", flush=True)
        net = PCNet(seed=a.seed)
        losses = []
        for e in range(a.epochs):  # one epoch per call (the optimiser state stays in the network)
            losses += net.train(samples, epochs=1, lr=a.lr).tolist()
            print(f"  epoch {e + 1}/{a.epochs}: loss {losses[-1]:.4f}", flush=True)
        print("loss per epoch:", np.round(losses, 4).tolist(), flush=True)
        out = a.weights if a.weights != WEIGHTS else WEIGHTS_FCG
        net.save(out)
        print(f"saved {out}", flush=True)
    elif a.mode == "train":
        samples = make_training_set(a.nreal, (n, n, n), a.nsteps, a.every)
        print(f"{len(samples)} samples at {n}^3", flush=True)
        net = PCNet(seed=a.seed)
        losses = net.train(samples, epochs=a.epochs, lr=a.lr)
        print("loss per epoch:", np.round(losses, 4).tolist(), flush=True)
        net.save(a.weights)
        print(f"saved {a.weights}", flush=True)
    else:
        net = PCNet()
        # the stage inside flexible CG has its own weights unless others are named
        net.load(WEIGHTS_FCG if a.fcg and a.weights == WEIGHTS else a.weights)
        case = io.workload("plain-small", seed=a.seed + 1000, grid=(n, n, n), corr_len=4.0)
        r = compare(case, net, a.nsteps, extrapolate=not a.no_extrapolate,
                    solver="cg" if a.fcg else "mixed", fcg=a.fcg, period=a.period)
        r["grid"], r["extrapolate"], r["stage"] = n, not a.no_extrapolate, "fcg" if a.fcg else "first"
        p0, p1 = r["plain"].pop("pressure"), r["net"].pop("pressure")
        r["p_rel_diff"] = float(np.abs(p1 - p0).max() / np.abs(p0).max())
        print(json.dumps(r), flush=True)
    net.close()


if __name__ == "__main__":
    main()
