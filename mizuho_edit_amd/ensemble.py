"""Ensembles: many WPFAW3D grids of different sizes on one GPU, stepped
together by the library's batched launches and captured graphs
(wpf_ensemble_* in include/wpf.h).  Not a reference feature: the reference
runs one realisation per job.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import WpfEnsembleOpts, WpfEnsembleStats, check
from .wpfaw3d import WPFAW3D


class Ensemble:
    """Members keep their own contexts; while the ensemble steps them their
    host-side time-loop state is stale until sync() (called by close() if
    needed).  After sync() the members may be stepped alone; the ensemble
    itself is then done with (make a new one to batch them again)."""

    def __init__(self, members: Sequence[WPFAW3D], opts, stages: int = 0, inner_iters: int = 0,
                 grid_blocks: int = 0, tile_depth: int = 0):
        self.members: List[WPFAW3D] = list(members)
        n = len(self.members)
        ptrs = (ctypes.c_void_p * n)(*[m._ctx.value for m in self.members])
        itime = (ctypes.c_int32 * n)(*[m.itime for m in self.members])
        time = np.array([m.time for m in self.members], dtype=np.float64)
        dtime = np.array([m.dtime for m in self.members], dtype=np.float64)
        self.opts = opts
        eo = WpfEnsembleOpts(stages, inner_iters, grid_blocks, tile_depth)
        h = ctypes.c_void_p()
        check(_lib.lib().wpf_ensemble_create(ptrs, n, ctypes.byref(opts), ctypes.byref(eo), itime,
                                             time.ctypes.data_as(_lib.D), dtime.ctypes.data_as(_lib.D),
                                             ctypes.byref(h)), "wpf_ensemble_create")
        self._h = h
        self._synced = True
        self.cells = sum(m.param.nx * m.param.ny * m.param.nz for m in self.members)

    def tune(self):
        g, z = ctypes.c_int32(), ctypes.c_int32()
        check(_lib.lib().wpf_ensemble_tune(self._h, ctypes.byref(g), ctypes.byref(z)), "wpf_ensemble_tune")
        return g.value, z.value

    def step(self, nsteps: int = 1) -> dict:
        st = WpfEnsembleStats()
        self._synced = False
        check(_lib.lib().wpf_ensemble_step(self._h, nsteps, ctypes.byref(st)), "wpf_ensemble_step")
        return st.as_dict()

    def sync(self):
        """Hand the members back: their time-loop state from the device."""
        n = len(self.members)
        itime = (ctypes.c_int32 * n)()
        time, dtime, taumin = (np.zeros(n) for _ in range(3))
        check(_lib.lib().wpf_ensemble_sync(self._h, itime, time.ctypes.data_as(_lib.D),
                                           dtime.ctypes.data_as(_lib.D), taumin.ctypes.data_as(_lib.D)),
              "wpf_ensemble_sync")
        for k, m in enumerate(self.members):
            m.itime, m.time, m.dtime, m.taumin = int(itime[k]), float(time[k]), float(dtime[k]), float(taumin[k])
        self._synced = True

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            if not self._synced:
                self.sync()
            _lib.lib().wpf_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
