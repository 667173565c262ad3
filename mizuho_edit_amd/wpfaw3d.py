"""Host-side mirror of ``module wPFAW3D`` (wPFAW3D.f90:19-2168) over the C ABI.

The class keeps the module's vocabulary: ``param``, ``well``, ``lw`` and the
field arrays are attributes, and the procedures of the per-time-step path
keep their names and their (absence of) arguments:

    calcPermeability()            wPFAW3D.f90:561-599
    solvePressure()               wPFAW3D.f90:603-1160
    calcVelocity()                wPFAW3D.f90:1164-1193
    calcPhaseDeriv() -> taumin    wPFAW3D.f90:1197-1475
    calcConcentDeriv()            wPFAW3D.f90:1554-1624
    evolvePhase(), evolveConcent() wPFAW3D.f90:1628-1640
    updateHalo(field)             wPFAW3D.f90:1645-1657
    getPressure(ix, iy, iz)       wPFAW3D.f90:2151-2166

The fields live in HBM; reading ``pressure``, ``phase``, ... downloads a
haloed Fortran-order copy.  A non-zero status raises ``WpfError``, as the
reference stops on a KSPSolve error (wPFAW3D.f90:1005).
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import _lib
from ._lib import FIELD, SOLVER, WpfError, WpfSolveStats, WpfSolverOpts, check
from .params import WpfDiag, wells_to_c


def solver_opts(solver: str = "mgcg", rtol: float = 1e-10, maxit: int = 1000, mg_levels: int = 5,
                mg_sweeps: int = 2, mg_omega: float = 0.8, mg_coarse_scale: float = 2.0,
                inner_rtol: float = 1e-4, inner_maxit: int = 100, flags: int = 0,
                atol: float = 1e-50) -> WpfSolverOpts:
    return WpfSolverOpts(solver=SOLVER[solver], maxit=maxit, rtol=rtol, atol=atol,
                         mg_levels=mg_levels, mg_sweeps=mg_sweeps, mg_omega=mg_omega,
                         mg_coarse_scale=mg_coarse_scale, inner_rtol=inner_rtol,
                         inner_maxit=inner_maxit, flags=flags)


class WPFAW3D:
    """One grid on one GPU, initialised from a loaded case (the state
    loadConfig + loadField leave behind)."""

    def __init__(self, case, device: int = -1, opts: Optional[WpfSolverOpts] = None):
        L = _lib.lib()
        self.case = case
        self.param = case.params
        self.well = case.wells
        self.lw = np.asfortranarray(case.lw, dtype=np.int32)
        self._cp = self.param.to_c()
        self._cw = wells_to_c(self.well)
        ctx = ctypes.c_void_p()
        check(L.wpf_create(ctypes.byref(self._cp),
                           self.lw.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                           self._cw, len(self.well), device, ctypes.byref(ctx)), "wpf_create")
        self._ctx = ctx
        self.opts = opts if opts is not None else solver_opts()
        self.stats = WpfSolveStats()
        self.upload("permeab_rock", case.krock)
        self.upload("phase", case.phase)
        self.upload("concent", case.concent)
        # time-loop state of program main (wPFAW3D.f90:2215-2218)
        self.itime = 0
        self.time = case.time_start
        self.dtime = self.param.dtime
        self.taumin = 0.0
        self.history = []

    # -- lifecycle ---------------------------------------------------------
    def close(self):
        if getattr(self, "_ctx", None) is not None and self._ctx.value:
            _lib.lib().wpf_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- host <-> device -----------------------------------------------------
    def _shape(self, comps: int = 0):
        p = self.param
        sh = (p.nx + 2, p.ny + 2, p.nz + 2)
        return ((comps,) + sh) if comps else sh

    def upload(self, name: str, arr: np.ndarray):
        a = np.asfortranarray(arr, dtype=np.float64)
        assert a.shape == self._shape(), (a.shape, self._shape())
        check(_lib.lib().wpf_upload(self._ctx, FIELD[name], a.ctypes.data_as(_lib.D)),
              f"wpf_upload({name})")

    def download(self, name: str) -> np.ndarray:
        comps = {"permeab": 4, "velocity": 3, "face_t": 3, "face_v": 3}.get(name, 0)
        out = np.zeros(self._shape(comps), dtype=np.float64, order="F")
        check(_lib.lib().wpf_download(self._ctx, FIELD[name], out.ctypes.data_as(_lib.D)),
              f"wpf_download({name})")
        return out

    permeab_rock = property(lambda s: s.download("permeab_rock"))
    pressure = property(lambda s: s.download("pressure"))
    phase = property(lambda s: s.download("phase"))
    concent = property(lambda s: s.download("concent"))
    dphasedt = property(lambda s: s.download("dphasedt"))
    dconcentdt = property(lambda s: s.download("dconcentdt"))
    permeab = property(lambda s: s.download("permeab"))
    velocity = property(lambda s: s.download("velocity"))

    # -- procedures of the hot path ------------------------------------------
    def calcPermeability(self):
        check(_lib.lib().wpf_calc_permeability(self._ctx), "calcPermeability")

    def solvePressure(self, opts: Optional[WpfSolverOpts] = None) -> dict:
        st = WpfSolveStats()
        o = opts if opts is not None else self.opts
        check(_lib.lib().wpf_solve_pressure(self._ctx, ctypes.byref(o), ctypes.byref(st)),
              "solvePressure")
        self.stats = st
        return st.as_dict()

    def calcVelocity(self):
        check(_lib.lib().wpf_calc_velocity(self._ctx), "calcVelocity")

    def calcPhaseDeriv(self) -> float:
        t = ctypes.c_double()
        check(_lib.lib().wpf_calc_phase_deriv(self._ctx, ctypes.byref(t)), "calcPhaseDeriv")
        return t.value

    def calcConcentDeriv(self):
        check(_lib.lib().wpf_calc_concent_deriv(self._ctx), "calcConcentDeriv")

    def evolvePhase(self):
        check(_lib.lib().wpf_evolve_phase(self._ctx, self.dtime), "evolvePhase")

    def evolveConcent(self):
        check(_lib.lib().wpf_evolve_concent(self._ctx, self.dtime), "evolveConcent")

    def updateHalo(self, name: str):
        check(_lib.lib().wpf_update_halo(self._ctx, FIELD[name]), "updateHalo")

    def getPressure(self, ix_g: int, iy_g: int, iz_g: int) -> float:
        v = ctypes.c_double()
        check(_lib.lib().wpf_get_pressure(self._ctx, ix_g, iy_g, iz_g, ctypes.byref(v)),
              "getPressure")
        return v.value

    def saveResult(self, filename: str, time: Optional[float] = None):
        """saveResult (wPFAW3D.f90:1842-2091): the first call writes the GiD
        mesh, every call appends Pressure, Phase-field, Concentration and
        Permeability (the node value) at `time` (default: the current time)."""
        first = not getattr(self, "_result_written", False)
        k0 = np.asfortranarray(self.permeab[0])
        arrs = [np.asfortranarray(a) for a in (self.pressure, self.phase, self.concent)] + [k0]
        check(_lib.lib().wpf_result_write(filename.encode(), self.time if time is None else time,
                                          ctypes.byref(self._cp), 1 if first else 0,
                                          *[a.ctypes.data_as(_lib.D) for a in arrs]), "saveResult")
        self._result_written = True

    def showProgress(self) -> WpfDiag:
        d = WpfDiag()
        check(_lib.lib().wpf_diagnostics(self._ctx, ctypes.byref(d)), "showProgress")
        return d

    # -- time loop of program main (wPFAW3D.f90:2225-2294) --------------------
    def step(self, opts: Optional[WpfSolverOpts] = None) -> dict:
        """One fused step: dt rule, permeability, pressure, fields."""
        o = opts if opts is not None else self.opts
        self.itime += 1
        t, dt, tm = ctypes.c_double(self.time), ctypes.c_double(self.dtime), ctypes.c_double(self.taumin)
        st = WpfSolveStats()
        # a solve that stops at maxit is reported, not fatal: the step is
        # completed, as the reference's loop continues after KSPSolve
        check(_lib.lib().wpf_step(self._ctx, ctypes.byref(o), self.itime, ctypes.byref(t),
                                  ctypes.byref(dt), ctypes.byref(tm), ctypes.byref(st)), "wpf_step",
              ok=(_lib.WPF_OK, _lib.WPF_ERR_NOCONV))
        self.time, self.dtime, self.taumin = t.value, dt.value, tm.value
        self.stats = st
        rec = st.as_dict()
        rec.update(itime=self.itime, time=self.time, dtime=self.dtime, taumin=self.taumin)
        self.history.append(rec)
        return rec

    def step_procedures(self, opts: Optional[WpfSolverOpts] = None) -> dict:
        """One step through the separate procedures, in main's order."""
        self.itime += 1
        if self.itime >= 3:
            p = self.param
            self.dtime = (p.tfac * 0.5 * (p.dx * p.dx) / (p.epsilon * p.epsilon)
                          * min((p.epsilon * p.epsilon) / p.diffusion, self.taumin))
        self.time = self.time + self.dtime
        self.calcPermeability()
        st = self.solvePressure(opts)
        self.calcVelocity()
        self.taumin = self.calcPhaseDeriv()
        self.calcConcentDeriv()
        self.evolvePhase()
        self.evolveConcent()
        st.update(itime=self.itime, time=self.time, dtime=self.dtime, taumin=self.taumin)
        self.history.append(st)
        return st

    def run(self, nsteps: int, opts: Optional[WpfSolverOpts] = None):
        for _ in range(nsteps):
            self.step(opts)
        return self

    # -- learned pressure correction -------------------------------------------
    def attach_net(self, net):
        """Use `net` (a learned.PCNet, kept alive by the caller) for solves
        whose flags ask for the learned stage; None detaches."""
        check(_lib.lib().wpf_attach_net(self._ctx, net.handle if net is not None else None),
              "attach_net")
        self._net = net

    def set_tile_depth(self, planes: int):
        """Planes per tile of the vector solver kernels: > 0 fixed, 0 the
        default, -1 autotuned on the next solve."""
        check(_lib.lib().wpf_set_tile_depth(self._ctx, planes), "set_tile_depth")

    # -- kernel timing ------------------------------------------------------
    def profile_start(self, kernel: str):
        check(_lib.lib().wpf_profile_start(self._ctx, _lib.KERNEL[kernel]), "profile_start")

    def profile_stop(self):
        ms, n = ctypes.c_double(), ctypes.c_int32()
        check(_lib.lib().wpf_profile_stop(self._ctx, ctypes.byref(ms), ctypes.byref(n)),
              "profile_stop")
        return ms.value, n.value

    # -- breakthrough (workload definition; not a reference quantity) --------
    def set_breakthrough_plane(self, ix_g: int):
        check(_lib.lib().wpf_set_breakthrough_plane(self._ctx, ix_g), "set_breakthrough_plane")

    def breakthrough_step(self) -> int:
        v = ctypes.c_int32()
        check(_lib.lib().wpf_breakthrough_step(self._ctx, ctypes.byref(v)), "breakthrough_step")
        return v.value
