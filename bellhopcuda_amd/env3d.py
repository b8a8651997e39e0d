"""Raw (user-unit) 3D environment descriptions and the 3D workload builders.

`Env3D` holds what a BELLHOP3D .env/.bty/.ati/.ssp file set holds for the
full-3D run (src/module/*.hpp Read() with O3D = true); `to_c()` packs it into
the `bhc_env3d` C struct of include/bhc_hip.h. Angle and receiver vectors are
given already expanded (see `env.subtab`).

Workload builders:
  analysis3d_env() -- the reference's isovelocity 3D arrivals test case.
  c3_env()       -- 3D Gaussian-beam TL over a 40 x 40 km shelf with a canyon,
                    hexahedral Munk-perturbed sound speed, 181 x 1001 beams,
                    36 bearings x 101 depths x 1001 ranges to 20 km.
  free_env3d()   -- isovelocity free space (closed-form spherical spreading).
  munk_env3d()   -- depth-only Munk profile over flat boundaries, any bearing
                    fan (the 3D path on an environment the 2D path also solves).
"""
import ctypes as C
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _abi
from .env import Halfspace, ReflTable, munk_c, subtab


@dataclass
class Grid3D:
    """Altimetry/bathymetry grid (x, y in km, z[ix, iy] in m); empty = flat default.
    type: 'R' (flat triangles) or 'C' (curvilinear), then 'S' or 'L'
    (provinces: province[ix, iy] 1-based numbers into prov rows of alphaR,
    betaR, rho, alphaI, betaI with attenuation in dB/wavelength)."""
    x_km: np.ndarray = field(default_factory=lambda: np.zeros(0))
    y_km: np.ndarray = field(default_factory=lambda: np.zeros(0))
    z: np.ndarray = field(default_factory=lambda: np.zeros((0, 0)))
    type: str = "RS"
    province: Optional[np.ndarray] = None
    prov: Optional[np.ndarray] = None


@dataclass
class Env3D:
    precision: int = 64
    freq0: float = 50.0
    ssp_type: str = "C"
    atten_unit: str = "W "
    ssp_z: np.ndarray = field(default_factory=lambda: np.zeros(0))
    ssp_c: np.ndarray = field(default_factory=lambda: np.zeros(0))
    ssp_alphaI: Optional[np.ndarray] = None
    ssp_rho: Optional[np.ndarray] = None
    hex_x_km: Optional[np.ndarray] = None     # 'H'
    hex_y_km: Optional[np.ndarray] = None
    hex_z: Optional[np.ndarray] = None
    hex_c: Optional[np.ndarray] = None        # [nx, ny, nz]
    fg: tuple = (20.0, 35.0, 8.0, 0.0)
    top: Halfspace = field(default_factory=lambda: Halfspace("V"))
    bot: Halfspace = field(default_factory=lambda: Halfspace("R"))
    ati: Grid3D = field(default_factory=Grid3D)
    bty: Grid3D = field(default_factory=Grid3D)
    trc: ReflTable = field(default_factory=ReflTable)
    brc: ReflTable = field(default_factory=ReflTable)
    Sx_km: np.ndarray = field(default_factory=lambda: np.zeros(1))
    Sy_km: np.ndarray = field(default_factory=lambda: np.zeros(1))
    Sz: np.ndarray = field(default_factory=lambda: np.zeros(1, np.float32))
    Rz: np.ndarray = field(default_factory=lambda: np.zeros(1, np.float32))
    Rr_km: np.ndarray = field(default_factory=lambda: np.zeros(1))
    theta_deg: np.ndarray = field(default_factory=lambda: np.zeros(1))
    alpha_deg: np.ndarray = field(default_factory=lambda: np.zeros(1))
    beta_deg: np.ndarray = field(default_factory=lambda: np.zeros(1))
    iSingle_alpha: int = 0
    iSingle_beta: int = 0
    run_type: str = "CB   3"
    beam_type: str = "   "
    deltas: float = 0.0
    box_x_km: float = 101.0
    box_y_km: float = 101.0
    box_z: float = 5500.0
    sbp: Optional[np.ndarray] = None
    bio: Optional[np.ndarray] = None          # [n, 5] z1, z2 (m), f0 (Hz), Q, a0 (dB/km); atten_unit[1] == 'B'

    @property
    def nx2d(self):
        """Nx2D run (RunType[5] == '2'): 2D rays in the vertical plane of each bearing."""
        return (self.run_type + " " * 6)[5] == "2"

    @property
    def nrays(self):
        na = 1 if self.iSingle_alpha > 0 else len(self.alpha_deg)
        # an Nx2D run that is not a ray run traces on the receiver bearings
        # (src/module/rayangles.hpp:163-178)
        fan = self.theta_deg if (self.nx2d and self.run_type[0] != "R") else self.beta_deg
        nb = 1 if self.iSingle_beta > 0 else len(fan)
        return na * nb * len(self.Sx_km) * len(self.Sy_km) * len(self.Sz)

    @property
    def field_shape(self):
        return (len(self.Sz), len(self.Sx_km), len(self.Sy_km), len(self.theta_deg), len(self.Rz),
                len(self.Rr_km))

    def to_c(self):
        """Return (Env3D_C, keepalive) -- the struct points into numpy arrays."""
        keep = []

        def dp(a):
            if a is None:
                return C.cast(None, _abi.c_double_p)
            arr = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
            keep.append(arr)
            return arr.ctypes.data_as(_abi.c_double_p)

        def fp(a):
            arr = np.ascontiguousarray(a, dtype=np.float32)
            keep.append(arr)
            return arr.ctypes.data_as(_abi.c_float_p)

        e = _abi.Env3D_C()
        e.precision = self.precision
        e.freq0 = self.freq0
        e.ssp_type = self.ssp_type.encode()
        e.atten_unit = (self.atten_unit + "  ")[:2].encode()
        n = len(self.ssp_z)
        e.ssp_npts = n
        e.ssp_z = dp(self.ssp_z)
        e.ssp_alphaR = dp(self.ssp_c)
        e.ssp_alphaI = dp(self.ssp_alphaI if self.ssp_alphaI is not None else np.zeros(n))
        e.ssp_rho = dp(self.ssp_rho if self.ssp_rho is not None else np.ones(n))
        e.ssp_betaR = dp(np.zeros(n))
        e.ssp_betaI = dp(np.zeros(n))
        if self.ssp_type == "H":
            e.hex_nx, e.hex_ny, e.hex_nz = len(self.hex_x_km), len(self.hex_y_km), len(self.hex_z)
            e.hex_x_km, e.hex_y_km, e.hex_z = dp(self.hex_x_km), dp(self.hex_y_km), dp(self.hex_z)
            e.hex_c = dp(np.asarray(self.hex_c, np.float64).reshape(-1))
        e.fg_T, e.fg_S, e.fg_pH, e.fg_z_bar = self.fg
        for dst, hs in ((e.top, self.top), (e.bot, self.bot)):
            dst.bc = hs.bc.encode()
            dst.depth, dst.alphaR, dst.betaR, dst.rho = hs.depth, hs.alphaR, hs.betaR, hs.rho
            dst.alphaI, dst.betaI, dst.zTemp, dst.Mz = hs.alphaI, hs.betaI, hs.zTemp, hs.Mz
        for dst, g in ((e.ati, self.ati), (e.bty, self.bty)):
            dst.type = (g.type + "  ")[:2].encode()
            dst.nx, dst.ny = len(g.x_km), len(g.y_km)
            if dst.nx:
                dst.x_km, dst.y_km, dst.z = dp(g.x_km), dp(g.y_km), dp(np.asarray(g.z, np.float64))
            if g.province is not None:
                pv = np.ascontiguousarray(g.province, dtype=np.int32).reshape(-1)
                keep.append(pv)
                dst.province = pv.ctypes.data_as(_abi.c_int32_p)
                pr = np.asarray(g.prov, np.float64).reshape(-1, 5)
                dst.nprov = len(pr)
                dst.prov = dp(pr)
        for dst, t in ((e.trc, self.trc), (e.brc, self.brc)):
            dst.npts = len(t.theta)
            if dst.npts:
                dst.theta, dst.r, dst.phi = dp(t.theta), dp(t.r), dp(t.phi)
        e.NSx, e.NSy, e.NSz = len(self.Sx_km), len(self.Sy_km), len(self.Sz)
        e.NRz, e.NRr, e.Ntheta = len(self.Rz), len(self.Rr_km), len(self.theta_deg)
        e.Sx_km, e.Sy_km, e.Sz = dp(self.Sx_km), dp(self.Sy_km), fp(self.Sz)
        e.Rz, e.Rr_km, e.theta_deg = fp(self.Rz), dp(self.Rr_km), dp(self.theta_deg)
        e.Nalpha, e.iSingle_alpha = len(self.alpha_deg), self.iSingle_alpha
        e.Nbeta, e.iSingle_beta = len(self.beta_deg), self.iSingle_beta
        e.alpha_deg, e.beta_deg = dp(self.alpha_deg), dp(self.beta_deg)
        e.RunType = (self.run_type + " " * 7)[:7].encode()
        e.BeamType = (self.beam_type + "   ")[:3].encode()
        e.deltas, e.box_x_km, e.box_y_km, e.box_z = self.deltas, self.box_x_km, self.box_y_km, self.box_z
        if self.sbp is not None:
            e.NSBPPts = len(self.sbp)
            e.sbp = dp(np.asarray(self.sbp, np.float64).reshape(-1))
        if self.bio is not None:
            e.NBioLayers = len(self.bio)
            e.bio = dp(np.asarray(self.bio, np.float64).reshape(-1))
        return e, keep


def bearings(n, dtype=np.float64):
    """n receiver bearings over a full circle, 360/n apart (the last of a
    closed 0..360 sweep removed, as the reference's CheckFix360Sweep does)."""
    return subtab(0.0, 360.0, n + 1, dtype)[:n]


def free_env3d(nalpha=91, nbeta=37, ntheta=4, nrz=21, nrr=51, precision=64, run_type="CB   3",
               r_max_km=5.0, sd=2500.0, freq=100.0):
    """Isovelocity (1500 m/s) free space, 0-5000 m, bounded above and below by
    halfspaces of the same medium (reflection coefficient 0: a ray that meets
    them ends); beams over +-89 deg declination and a full circle of bearings."""
    z = np.array([0.0, 5000.0])
    matched = dict(alphaR=1500.0, betaR=0.0, rho=1.0, alphaI=0.0)
    return Env3D(
        precision=precision, freq0=freq, ssp_type="C", ssp_z=z, ssp_c=np.array([1500.0, 1500.0]),
        top=Halfspace("A", depth=0.0, zTemp=0.0, **matched), bot=Halfspace("A", depth=5000.0, zTemp=5000.0, **matched),
        Sz=np.array([sd], np.float32), Rz=subtab(sd - 1000.0, sd + 1000.0, nrz, np.float32),
        Rr_km=subtab(0.0, r_max_km, nrr), theta_deg=bearings(ntheta),
        alpha_deg=subtab(-89.0, 89.0, nalpha), beta_deg=bearings(nbeta) - 180.0,
        run_type=run_type, deltas=100.0, box_x_km=r_max_km + 1, box_y_km=r_max_km + 1, box_z=5001.0)


def analysis3d_env(precision=64, run_type="AG   3"):
    """The reference's 3D arrivals test case (test/individual/analysis_3d/
    analysis_3d.env): isovelocity 1000 m/s water 0-1500 m, vacuum surface,
    near-matched halfspace bottom; 4000 Hz source at (0, 0, 100 m); receivers
    at 100 m depth, 0.1-1.5 km on the 34 deg bearing; 3 declinations on
    [-1, 1] deg x bearings 34 and 35 deg; geometric hat beams, deltas 1.7646 m,
    box 1.65 km x 1.65 km x 1650 m."""
    z = np.array([0.0, 100.0, 500.0, 1500.0])
    return Env3D(
        precision=precision, freq0=4000.0, ssp_type="C", atten_unit="F ", ssp_z=z, ssp_c=np.full(4, 1000.0),
        top=Halfspace("V", depth=0.0),
        bot=Halfspace("A", depth=1500.0, alphaR=1000.0, betaR=0.0, rho=0.9, alphaI=0.0, zTemp=1000.0),
        Sz=np.array([100.0], np.float32), Rz=np.array([100.0], np.float32),
        Rr_km=subtab(0.1, 1.5, 15), theta_deg=np.array([34.0]),
        alpha_deg=subtab(-1.0, 1.0, 3), beta_deg=np.array([34.0, 35.0]),
        run_type=run_type, deltas=1.7646, box_x_km=1.65, box_y_km=1.65, box_z=1650.0)


def munk_env3d(nalpha=201, nbeta=5, beta=(-2.0, 2.0), ntheta=1, theta0=0.0, nrz=51, nrr=201, precision=64,
               run_type="CB   3", r_max_km=20.0, freq=50.0, sd=1000.0, ssp_type="C", bot="A"):
    """Depth-only Munk profile (51 knots, 0-5000 m) over flat boundaries: vacuum
    surface, acousto-elastic bottom; receivers on bearing(s) theta."""
    z = np.linspace(0.0, 5000.0, 51)
    th = np.array([theta0]) if ntheta == 1 else bearings(ntheta)
    return Env3D(
        precision=precision, freq0=freq, ssp_type=ssp_type, ssp_z=z, ssp_c=munk_c(z),
        top=Halfspace("V", depth=0.0),
        bot=Halfspace(bot, depth=5000.0, alphaR=1600.0, betaR=0.0, rho=1.8, alphaI=0.8, zTemp=5000.0),
        Sz=np.array([sd], np.float32), Rz=subtab(0.0, 5000.0, nrz, np.float32),
        Rr_km=subtab(0.0, r_max_km, nrr), theta_deg=th,
        alpha_deg=subtab(-20.0, 20.0, nalpha), beta_deg=subtab(beta[0], beta[1], nbeta),
        run_type=run_type, deltas=0.0, box_x_km=r_max_km + 1, box_y_km=r_max_km + 1, box_z=5500.0)


def c3_env(seed=20260003, nalpha=1001, nbeta=181, ntheta=36, nrz=101, nrr=1001, precision=64,
           run_type="CB   3", nxy=201, hex_n=(21, 21, 50)):
    """3D Gaussian-beam TL: 201 x 201 bathymetry grid over 40 x 40 km (source at
    the centre), a shelf sloping from 200 m to 2000 m along x with a 300 m deep
    Gaussian canyon along x; hexahedral 21 x 21 x 50 sound speed from a seeded
    smooth 3D perturbation (+-10 m/s) of the Munk profile; source 100 m, 200 Hz;
    nbeta bearings x nalpha declinations; receivers ntheta bearings x nrz depths
    x nrr ranges to 20 km."""
    rng = np.random.default_rng(seed)
    x = np.linspace(-20.0, 20.0, nxy)
    y = np.linspace(-20.0, 20.0, nxy)
    X, Y = np.meshgrid(x, y, indexing="ij")
    shelf = 200.0 + 1800.0 * (X + 20.0) / 40.0
    canyon = 300.0 * np.exp(-0.5 * (Y / 2.0) ** 2)
    depth = shelf + canyon
    zmax = float(depth.max()) + 100.0
    nx, ny, nz = hex_n
    hx = np.linspace(-22.0, 22.0, nx)
    hy = np.linspace(-22.0, 22.0, ny)
    hz = np.linspace(0.0, zmax, nz)
    pert = np.zeros((nx, ny, nz))
    ux, uy, uz = np.linspace(0, 1, nx), np.linspace(0, 1, ny), np.linspace(0, 1, nz)
    for _ in range(6):
        kx, ky, kz = rng.integers(1, 4, 3)
        ph = rng.uniform(0, 2 * np.pi, 3)
        pert += rng.normal() * (np.sin(np.pi * kx * ux + ph[0])[:, None, None]
                                * np.sin(np.pi * ky * uy + ph[1])[None, :, None]
                                * np.sin(np.pi * kz * uz + ph[2])[None, None, :])
    pert *= 10.0 / max(np.max(np.abs(pert)), 1e-12)
    cgrid = munk_c(hz)[None, None, :] + pert
    z1 = hz.copy()
    return Env3D(
        precision=precision, freq0=200.0, ssp_type="H", ssp_z=z1, ssp_c=munk_c(z1),
        hex_x_km=hx, hex_y_km=hy, hex_z=hz, hex_c=cgrid,
        top=Halfspace("V", depth=0.0),
        bot=Halfspace("A", depth=zmax, alphaR=1700.0, betaR=0.0, rho=1.8, alphaI=0.5, zTemp=zmax),
        bty=Grid3D(x_km=x, y_km=y, z=depth),
        Sx_km=np.zeros(1), Sy_km=np.zeros(1), Sz=np.array([100.0], np.float32),
        Rz=subtab(0.0, 2000.0, nrz, np.float32), Rr_km=subtab(0.0, 20.0, nrr),
        theta_deg=bearings(ntheta),
        alpha_deg=subtab(-60.0, 60.0, nalpha), beta_deg=bearings(nbeta) - 180.0,
        run_type=run_type, deltas=0.0, box_x_km=19.9, box_y_km=19.9, box_z=zmax + 50.0)
