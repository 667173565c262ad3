"""Input pipeline of wPFAW3D: ``.cfg`` + well file + ``wPFAWHFG`` field file,
and the seeded synthetic rock realisations the workloads run on.

* ``write_config`` / ``read_config`` follow loadConfig (wPFAW3D.f90:264-505):
  ``key = value`` lines, ``#`` comments, 24 keys, derived ``dx`` and
  ``alpha`` (1:390-399), a well file whose 2nd line is the well count
  (1:401-432).
* ``write_field`` / ``read_field`` follow saveField / loadField
  (wPFAW3D.f90:2096-2141, 509-557): 32-byte header (magic ``wPFAWHFG``,
  version ``20250304``, time, nx/ny/nz as int16, pad), then permeab_rock,
  phase, concent as FP64 global arrays in Fortran order.
* ``gaussian_random_field`` is a seeded, portable GRF: splitmix64 counters
  and Box-Muller white noise convolved with a separable Gaussian kernel,
  normalised by the kernel energy, covariance exp(-r^2/(2 l^2)).
"""
from __future__ import annotations

import math
import os
import struct
from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np

from .params import Params, Well, well_index_table

FIELD_MAGIC = b"wPFAWHFG"
FIELD_VERSION = b"20250304"
# The smallest normal double: a "p = 0" outlet that the reference still
# treats as Dirichlet, since onBoundary requires pressure /= 0
# (wPFAW3D.f90:1106).
OUTLET_SENTINEL = 2.2250738585072014e-308


# --------------------------------------------------------------------------
# .cfg + well file
# --------------------------------------------------------------------------
_CFG_KEYS = ("grid", "dt", "tfac", "time_total", "gid_output", "scr_output", "w", "dxw",
             "lambda", "poro", "rhos", "rhoa", "visco", "cwtp", "D", "b100", "ral", "rat",
             "kwh", "pini", "ref", "well", "init", "last")


def _fmt(x: float) -> str:
    return repr(float(x))


def write_config(path: str, p: Params, well_file: str, init_file: Optional[str] = None,
                 last_file: str = "last.fld") -> None:
    lines = [
        "# wPFAW3D configuration (keys of loadConfig, wPFAW3D.f90:305-385)",
        f"grid = {p.nx} {p.ny} {p.nz}",
        f"dt = {_fmt(p.dtime)}",
        f"tfac = {_fmt(p.tfac)}",
        f"time_total = {_fmt(p.time_total)}",
        f"gid_output = {p.n_res_out}",
        f"scr_output = {p.n_dat_out}",
        f"w = {_fmt(p.epsilon)}",
        f"dxw = {_fmt(p.dxw)}",
        f"lambda = {_fmt(p.lambda_)}",
        f"poro = {_fmt(p.porosity)}",
        f"rhos = {_fmt(p.rhos)}",
        f"rhoa = {_fmt(p.rhoa)}",
        f"visco = {_fmt(p.viscosity)}",
        f"cwtp = {_fmt(p.cwtp)}",
        f"D = {_fmt(p.diffusion)}",
        f"b100 = {_fmt(p.b100)}",
        f"ral = {_fmt(p.ral)}",
        f"rat = {_fmt(p.rat)}",
        f"kwh = {_fmt(p.permeab_hole)}",
        f"pini = {_fmt(p.pini)}",
        f"ref = {p.ix_ref} {p.iy_ref} {p.iz_ref}",
        f"well = {well_file}",
    ]
    if init_file:
        lines.append(f"init = {init_file}")
    lines.append(f"last = {last_file}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_wells(path: str, wells: List[Well]) -> None:
    with open(path, "w") as f:
        f.write("# ix iy iz wpq value phase concent\n")
        f.write(f"{len(wells)}\n")
        for w in wells:
            f.write(f"{w.ix_g} {w.iy_g} {w.iz_g} {w.wpq} {_fmt(w.wval)} "
                    f"{_fmt(w.phase)} {_fmt(w.concent)}\n")


class ConfigError(ValueError):
    """Raised where loadConfig calls exit(1)."""


def _split_values(s: str) -> List[str]:
    return s.replace(",", " ").split()


def read_config(path: str):
    """Parse a ``.cfg`` the way loadConfig does.  Returns (Params, wells,
    init) where ``init`` is either a field-file path or a tuple
    (init_permeab, init_phase, init_concent)."""
    if not os.path.exists(path):
        raise ConfigError(f"can not open config file: {path}")
    p = Params()
    p.dtime = 0.1e-6  # default, 1:286
    init: object = (0.0, -1.0, 0.0)
    well_file = None
    base = os.path.dirname(os.path.abspath(path))
    with open(path) as f:
        for raw in f:
            buf = raw.rstrip("\n")
            if buf.startswith("#"):
                continue
            i = buf.find("#")
            if i >= 0:
                buf = buf[:i]
            if not buf.strip():
                continue
            i = buf.find("=")
            if i >= 0:
                buf = buf[:i] + " " + buf[i + 1:]
            tok = _split_values(buf)
            key, vals = tok[0], tok[1:]
            try:
                if key == "grid":
                    p.nx, p.ny, p.nz = (int(v) for v in vals[:3])
                elif key == "dt":
                    p.dtime = float(vals[0])
                elif key == "tfac":
                    p.tfac = float(vals[0])
                elif key == "time_total":
                    p.time_total = float(vals[0])
                elif key == "gid_output":
                    p.n_res_out = int(vals[0])
                elif key == "scr_output":
                    p.n_dat_out = int(vals[0])
                elif key == "w":
                    p.epsilon = float(vals[0])
                elif key == "dxw":
                    p.dxw = float(vals[0])
                elif key == "lambda":
                    p.lambda_ = float(vals[0])
                elif key == "poro":
                    p.porosity = float(vals[0])
                elif key == "rhos":
                    p.rhos = float(vals[0])
                elif key == "rhoa":
                    p.rhoa = float(vals[0])
                elif key == "visco":
                    p.viscosity = float(vals[0])
                elif key == "cwtp":
                    p.cwtp = float(vals[0])
                elif key == "D":
                    p.diffusion = float(vals[0])
                elif key == "b100":
                    p.b100 = float(vals[0])
                elif key == "ral":
                    p.ral = float(vals[0])
                elif key == "rat":
                    p.rat = float(vals[0])
                elif key == "kwh":
                    p.permeab_hole = float(vals[0])
                elif key == "pini":
                    p.pini = float(vals[0])
                elif key == "ref":
                    p.ix_ref, p.iy_ref, p.iz_ref = (int(v) for v in vals[:3])
                elif key == "well":
                    well_file = vals[0]
                elif key == "init":
                    try:
                        init = tuple(float(v) for v in vals[:3])
                        if len(init) != 3:
                            raise ValueError
                    except ValueError:
                        init = vals[0]
                elif key == "last":
                    pass
                else:
                    raise ConfigError(f"unknown key: {key}")
            except (IndexError, ValueError) as e:
                raise ConfigError(f"broken line: {raw.strip()}") from e
    if well_file is None:
        raise ConfigError("can not open well file: (none)")
    wpath = well_file if os.path.isabs(well_file) else os.path.join(base, well_file)
    wells = read_wells(wpath)
    if isinstance(init, str) and not os.path.isabs(init):
        init = os.path.join(base, init)
    return p, wells, init


def read_wells(path: str) -> List[Well]:
    if not os.path.exists(path):
        raise ConfigError(f"can not open well file: {path}")
    with open(path) as f:
        f.readline()
        nwell = int(_split_values(f.readline())[0])
        wells = []
        for _ in range(nwell):
            t = _split_values(f.readline())
            wpq = int(t[3])
            if wpq not in (1, -1):
                raise ConfigError(f"invalid data in well file: {path}")
            wells.append(Well(int(t[0]), int(t[1]), int(t[2]), wpq, float(t[4]),
                              float(t[5]), float(t[6])))
    return wells


# --------------------------------------------------------------------------
# wPFAWHFG field file
# --------------------------------------------------------------------------
_HEADER = struct.Struct("<8s8sdhhhh")  # 32 bytes


def write_field(path: str, time: float, krock: np.ndarray, phase: np.ndarray,
                concent: np.ndarray) -> None:
    """Bodies are interior (global) arrays, Fortran order."""
    nx, ny, nz = krock.shape
    with open(path, "wb") as f:
        f.write(_HEADER.pack(FIELD_MAGIC, FIELD_VERSION, float(time), nx, ny, nz, 0))
        for a in (krock, phase, concent):
            f.write(np.asarray(a, dtype="<f8").tobytes(order="F"))


def read_field(path: str, grid: Optional[Tuple[int, int, int]] = None):
    if not os.path.exists(path):
        raise ConfigError(f"can not open field file: {path}")
    with open(path, "rb") as f:
        magic, _ver, time, nx, ny, nz, _ = _HEADER.unpack(f.read(_HEADER.size))
        if magic != FIELD_MAGIC:
            raise ConfigError(f"invalid magic in field file: {path}")
        if grid is not None and (nx, ny, nz) != tuple(grid):
            raise ConfigError(f"invalid grid in field file: {path}")
        n = nx * ny * nz
        out = []
        for _ in range(3):
            a = np.frombuffer(f.read(8 * n), dtype="<f8")
            if a.size != n:
                raise ConfigError(f"short field file: {path}")
            out.append(a.reshape((nx, ny, nz), order="F").astype(np.float64))
    return time, out[0], out[1], out[2]


# --------------------------------------------------------------------------
# seeded Gaussian random field
# --------------------------------------------------------------------------
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def white_noise(shape: Tuple[int, ...], seed: int) -> np.ndarray:
    """Standard normal samples from counter-based splitmix64 + Box-Muller."""
    n = int(np.prod(shape))
    key = _splitmix64(np.array([seed], dtype=np.uint64))[0]
    idx = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        a = _splitmix64(key ^ (idx * np.uint64(2)))
        b = _splitmix64(key ^ (idx * np.uint64(2) + np.uint64(1)))
    u1 = ((a >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)
    u2 = ((b >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)
    z = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)
    return z.reshape(shape, order="F")


def gaussian_kernel(corr_len: float) -> np.ndarray:
    sig = corr_len / math.sqrt(2.0)
    r = int(math.ceil(4.0 * corr_len))
    x = np.arange(-r, r + 1, dtype=np.float64)
    return np.exp(-0.5 * (x / sig) ** 2)


def gaussian_random_field(shape: Tuple[int, int, int], corr_len: float, seed: int,
                          variance: float = 1.0) -> np.ndarray:
    """Zero-mean GRF with covariance variance*exp(-r^2/(2 corr_len^2))."""
    from scipy.ndimage import convolve1d

    g = gaussian_kernel(corr_len)
    r = (g.size - 1) // 2
    pad = tuple(s + 2 * r for s in shape)
    a = white_noise(pad, seed)
    for ax in range(3):
        a = convolve1d(a, g, axis=ax, mode="constant")
    a = a[r:r + shape[0], r:r + shape[1], r:r + shape[2]]
    a = a / (np.sum(g * g) ** 1.5)
    return np.asfortranarray(a * math.sqrt(variance))


def kozeny_carman(phi: np.ndarray, k0: float = 1.0e-12) -> np.ndarray:
    return k0 * phi ** 3 / (1.0 - phi) ** 2


# --------------------------------------------------------------------------
# workloads
# --------------------------------------------------------------------------
@dataclass
class Case:
    """Everything loadConfig leaves in the module: parameters, wells, the
    well index table and the three haloed state fields."""
    params: Params
    wells: List[Well]
    lw: np.ndarray
    krock: np.ndarray
    phase: np.ndarray
    concent: np.ndarray
    time_start: float = 0.0

    def haloed_zeros(self) -> np.ndarray:
        p = self.params
        return np.zeros((p.nx + 2, p.ny + 2, p.nz + 2), dtype=np.float64, order="F")


def clamp_halo(a: np.ndarray) -> np.ndarray:
    """updateHalo on one rank (wPFAW3D.f90:1651-1656): x, then y, then z."""
    a[0, :, :] = a[1, :, :]
    a[-1, :, :] = a[-2, :, :]
    a[:, 0, :] = a[:, 1, :]
    a[:, -1, :] = a[:, -2, :]
    a[:, :, 0] = a[:, :, 1]
    a[:, :, -1] = a[:, :, -2]
    return a


def inlet_outlet_wells(p: Params, p_in: float = 1.0e5, c_in: float = -1.0) -> List[Well]:
    """Inlet plane ix_g = 1 at p_in with phase +1 and concentration c_in;
    outlet plane ix_g = nx at "p = 0" (the sentinel) with phase -1, c = 0."""
    wells = []
    for k in range(1, p.nz + 1):
        for j in range(1, p.ny + 1):
            wells.append(Well(1, j, k, 1, p_in, 1.0, c_in))
    for k in range(1, p.nz + 1):
        for j in range(1, p.ny + 1):
            wells.append(Well(p.nx, j, k, 1, OUTLET_SENTINEL, -1.0, 0.0))
    return wells


def build_case(p: Params, wells: List[Well], krock_interior: np.ndarray,
               init_phase: float = -1.0, init_concent: float = 0.0) -> Case:
    """State after loadConfig + loadField (wPFAW3D.f90:440-496, 509-557)."""
    sh = (p.nx + 2, p.ny + 2, p.nz + 2)
    lw = well_index_table(p, wells)
    krock = np.zeros(sh, order="F")
    phase = np.full(sh, init_phase, order="F")
    conc = np.full(sh, init_concent, order="F")
    krock[1:-1, 1:-1, 1:-1] = krock_interior
    for w in wells:
        if 1 <= w.ix_g <= p.nx and 1 <= w.iy_g <= p.ny and 1 <= w.iz_g <= p.nz:
            phase[w.ix_g, w.iy_g, w.iz_g] = w.phase
            conc[w.ix_g, w.iy_g, w.iz_g] = w.concent
    for a in (krock, phase, conc):
        clamp_halo(a)
    return Case(p, wells, lw, krock, phase, conc)


def rock_permeability(shape, corr_len: float = 4.0, seed: int = 0, mean: float = 0.2,
                      std: float = 0.05, clip=(0.05, 0.35), variance: float = 1.0) -> np.ndarray:
    g = gaussian_random_field(shape, corr_len, seed, variance)
    phi = np.clip(mean + std * g, clip[0], clip[1])
    return kozeny_carman(phi)


def physics_params(nx: int, ny: int, nz: int, **kw) -> Params:
    """The physical constants every workload shares (dx = 1e-4 m from
    w = 3e-4 m and dxw = 1/3; D = 1e-9; mu = 1e-3); the values the workload
    descriptions leave open are fixed here once."""
    p = Params(nx=nx, ny=ny, nz=nz)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def workload(name: str, seed: int = 1, grid=None, **overrides) -> Case:
    """Synthetic inputs of the named workloads.

    plain-small : 64x32x32, corr 4, var 1, phi clip [0.05, 0.35]
    plain-256   : 256^3, same rock statistics
    mixed-256   : 256^3, GRF variance 2, clip [0.02, 0.45]
    learned-192 : 192^3, rock statistics of plain-256
    """
    sizes = {"plain-small": (64, 32, 32), "plain-256": (256, 256, 256),
             "mixed-256": (256, 256, 256), "learned-192": (192, 192, 192)}
    nx, ny, nz = grid if grid is not None else sizes[name]
    corr = overrides.pop("corr_len", 4.0)
    mean = overrides.pop("phi_mean", 0.2)
    if name == "mixed-256":
        kr = rock_permeability((nx, ny, nz), corr, seed, mean, 0.05, (0.02, 0.45), variance=2.0)
    else:
        kr = rock_permeability((nx, ny, nz), corr, seed, mean)
    c_in = overrides.pop("c_in", -1.0)
    p_in = overrides.pop("p_in", 1.0e5)
    p = physics_params(nx, ny, nz, **overrides)
    wells = inlet_outlet_wells(p, p_in=p_in, c_in=c_in)
    return build_case(p, wells, kr)


# the ensemble workload: (count, grid) groups of uneven realisations
ENSEMBLE_GROUPS = [(16, (64, 64, 64)), (16, (96, 64, 64)), (8, (128, 96, 96)), (8, (160, 128, 128))]


def ensemble_cases(seed0: int = 1000, shrink: int = 1, groups=None) -> List[Case]:
    """Seeded realisations of the plain workload's physics, one per member of
    the ensemble groups (48 grids, 40,894,464 cells at full size), each with
    its own correlation length in [2, 8] cells and mean porosity in
    [0.12, 0.25] (a low-discrepancy spread over the members).  shrink divides
    every grid dimension (tests)."""
    groups = groups or ENSEMBLE_GROUPS
    total = sum(cnt for cnt, _ in groups)
    cases, m = [], 0
    for cnt, (nx, ny, nz) in groups:
        for _ in range(cnt):
            u = (m * 0.6180339887498949) % 1.0
            v = (m * 0.7548776662466927) % 1.0
            grid = (nx // shrink, ny // shrink, nz // shrink)
            cases.append(workload("plain-small", seed=seed0 + m, grid=grid, corr_len=2.0 + 6.0 * u,
                                  phi_mean=0.12 + 0.13 * v))
            m += 1
    assert m == total
    return cases
