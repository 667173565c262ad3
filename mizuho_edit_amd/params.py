"""Parameter and well records shared by the host mirror, the C ABI and tests.

``WpfParams`` and ``WpfWell`` are the ctypes images of ``wpf_params`` and
``wpf_well`` in ``include/wpf.h``; they carry the subset of the reference's
``Param_type`` (wPFAW3D.f90:27-76) and ``Well_type`` (wPFAW3D.f90:82-90) that
the per-time-step path reads.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List

import numpy as np


class WpfParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("ix_ref", ctypes.c_int32), ("iy_ref", ctypes.c_int32), ("iz_ref", ctypes.c_int32),
        ("dtime", ctypes.c_double), ("tfac", ctypes.c_double),
        ("epsilon", ctypes.c_double), ("dx", ctypes.c_double),
        ("lambda_", ctypes.c_double), ("porosity", ctypes.c_double),
        ("rhos", ctypes.c_double), ("rhoa", ctypes.c_double),
        ("viscosity", ctypes.c_double), ("cwtp", ctypes.c_double),
        ("diffusion", ctypes.c_double), ("b100", ctypes.c_double),
        ("ral", ctypes.c_double), ("rat", ctypes.c_double),
        ("permeab_hole", ctypes.c_double), ("alpha", ctypes.c_double),
        ("pini", ctypes.c_double),
    ]


class WpfWell(ctypes.Structure):
    _fields_ = [
        ("ix_g", ctypes.c_int32), ("iy_g", ctypes.c_int32), ("iz_g", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("pressure", ctypes.c_double), ("injection", ctypes.c_double),
        ("phase", ctypes.c_double), ("concent", ctypes.c_double),
    ]


@dataclass
class Params:
    """Configuration keys of loadConfig (wPFAW3D.f90:305-385) plus the
    derived quantities it computes (1:390-399)."""
    nx: int = 64
    ny: int = 32
    nz: int = 32
    dtime: float = 1.0e-7          # "dt", default of 1:286
    tfac: float = 0.4
    time_total: float = 1.0e30
    n_res_out: int = 1
    n_dat_out: int = 1
    epsilon: float = 3.0e-4        # "w"
    dxw: float = 1.0 / 3.0
    lambda_: float = 0.8
    porosity: float = 0.2
    rhos: float = 2710.0
    rhoa: float = 1140.0
    viscosity: float = 1.0e-3
    cwtp: float = 0.28
    diffusion: float = 1.0e-9
    b100: float = 1.37
    ral: float = 1.0
    rat: float = 0.1
    permeab_hole: float = 1.0e-11
    pini: float = 0.0
    ix_ref: int = 1
    iy_ref: int = 1
    iz_ref: int = 1

    @property
    def dx(self) -> float:
        # param%dx = dxw*param%epsilon  (wPFAW3D.f90:393)
        return self.dxw * self.epsilon

    @property
    def alpha(self) -> float:
        # b = rhoa/rhos; kc = cwtp*b100/(1-poro); alpha = 0.5/b/kc  (1:397-399)
        b = self.rhoa / self.rhos
        kc = self.cwtp * self.b100 / (1.0 - self.porosity)
        return 0.5 / b / kc

    @property
    def ncell(self) -> int:
        return self.nx * self.ny * self.nz

    def to_c(self) -> WpfParams:
        return WpfParams(
            nx=self.nx, ny=self.ny, nz=self.nz,
            ix_ref=self.ix_ref, iy_ref=self.iy_ref, iz_ref=self.iz_ref,
            dtime=self.dtime, tfac=self.tfac, epsilon=self.epsilon, dx=self.dx,
            lambda_=self.lambda_, porosity=self.porosity, rhos=self.rhos, rhoa=self.rhoa,
            viscosity=self.viscosity, cwtp=self.cwtp, diffusion=self.diffusion,
            b100=self.b100, ral=self.ral, rat=self.rat, permeab_hole=self.permeab_hole,
            alpha=self.alpha, pini=self.pini)


@dataclass
class Well:
    """One well grid point; ``wpq`` +1 = fixed pressure, -1 = injection
    (wPFAW3D.f90:416-430)."""
    ix_g: int
    iy_g: int
    iz_g: int
    wpq: int
    wval: float
    phase: float
    concent: float

    @property
    def pressure(self) -> float:
        return self.wval if self.wpq == 1 else 0.0

    @property
    def injection(self) -> float:
        return self.wval if self.wpq == -1 else 0.0


def wells_to_c(wells: List[Well]):
    arr = (WpfWell * max(1, len(wells)))()
    for i, w in enumerate(wells):
        arr[i] = WpfWell(ix_g=w.ix_g, iy_g=w.iy_g, iz_g=w.iz_g, pad=0,
                         pressure=w.pressure, injection=w.injection,
                         phase=w.phase, concent=w.concent)
    return arr


def well_index_table(p: Params, wells: List[Well]) -> np.ndarray:
    """``lw(0:nx+1,0:ny+1,0:nz+1)`` as loadConfig builds it on one rank
    (wPFAW3D.f90:451-461): 1-based well index, later wells overwrite earlier
    ones, halo positions included."""
    lw = np.zeros((p.nx + 2, p.ny + 2, p.nz + 2), dtype=np.int32, order="F")
    for iw, w in enumerate(wells, start=1):
        ix, iy, iz = w.ix_g, w.iy_g, w.iz_g
        if 0 <= ix <= p.nx + 1 and 0 <= iy <= p.ny + 1 and 0 <= iz <= p.nz + 1:
            lw[ix, iy, iz] = iw
    return lw


class WpfDiag(ctypes.Structure):
    """``wpf_diag``: the showProgress reductions (wPFAW3D.f90:1695-1810)."""
    _fields_ = [(n, ctypes.c_double) for n in
                ("pv", "vcal", "rwhmax", "phase_max", "phase_min", "concent_max",
                 "concent_min", "pinj")]
