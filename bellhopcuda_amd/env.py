"""Raw (user-unit) 2D environment descriptions and the workload builders.

`Env2D` holds what a BELLHOP .env/.bty/.ati/.brc file set holds (the reference
reads these in src/module/*.hpp Read()); `to_c()` packs it into the
`bhc_env2d` C struct of include/bhc_hip.h. Vectors that an .env file gives as
"first last /" are expanded with `subtab`, which reproduces the reference's
SubTab (src/common_setup.hpp:282-316: x[i] = fma(i, (x1-x0)/(N-1), x0) in the
vector's own precision).

Workload builders:
  munk_env()      -- Munk profile, flat boundaries, 2001 hat beams, 201 x 1001 grid.
  c2_env()        -- range-dependent spline SSP + seamount bathymetry, 20001
                     ray-centred Gaussian beams, 501 x 5001 grid.
  sweep_envs()    -- the 4096-problem heterogeneous batched sweep.
"""
import ctypes as C
import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _abi

_libm = C.CDLL("libm.so.6")
_libm.fma.restype = C.c_double
_libm.fma.argtypes = [C.c_double, C.c_double, C.c_double]
_libm.fmaf.restype = C.c_float
_libm.fmaf.argtypes = [C.c_float, C.c_float, C.c_float]


def subtab(x0, x1, n, dtype=np.float64):
    """SubTab (src/common_setup.hpp:282-316) in float64 or float32, through
    the library's bhc_subtab (one fused multiply-add per element)."""
    if n <= 0:
        return np.zeros(0, dtype)
    if n == 1:
        return np.array([x0], dtype)
    from . import lib
    out = np.zeros(n, np.float32 if dtype == np.float32 else np.float64)
    rc = lib().bhc_subtab(float(x0), float(x1), n, 32 if dtype == np.float32 else 64,
                          out.ctypes.data_as(C.c_void_p))
    if rc != 0:
        raise ValueError("subtab failed")
    return out


@dataclass
class Halfspace:
    bc: str = "V"          # 'V' vacuum, 'R' rigid, 'A' halfspace, 'F' table, 'G' grain size
    depth: float = 0.0
    alphaR: float = 0.0
    betaR: float = 0.0
    rho: float = 1.0
    alphaI: float = 0.0
    betaI: float = 0.0
    zTemp: float = 0.0
    Mz: float = 0.0


@dataclass
class Boundary:
    """Altimetry/bathymetry points (range km, depth m); empty = flat default."""
    type: str = "LS"
    r_km: np.ndarray = field(default_factory=lambda: np.zeros(0))
    z: np.ndarray = field(default_factory=lambda: np.zeros(0))
    alphaR: Optional[np.ndarray] = None
    betaR: Optional[np.ndarray] = None
    rho: Optional[np.ndarray] = None
    alphaI: Optional[np.ndarray] = None
    betaI: Optional[np.ndarray] = None


@dataclass
class ReflTable:
    theta: np.ndarray = field(default_factory=lambda: np.zeros(0))
    r: np.ndarray = field(default_factory=lambda: np.zeros(0))
    phi: np.ndarray = field(default_factory=lambda: np.zeros(0))


@dataclass
class Env2D:
    title: str = "bellhop-mi355x"
    precision: int = 64
    freq0: float = 50.0
    ssp_type: str = "C"
    atten_unit: str = "W "
    ssp_z: np.ndarray = field(default_factory=lambda: np.zeros(0))
    ssp_c: np.ndarray = field(default_factory=lambda: np.zeros(0))       # alphaR
    ssp_alphaI: Optional[np.ndarray] = None
    ssp_rho: Optional[np.ndarray] = None
    ssp_r_km: Optional[np.ndarray] = None     # 'Q'
    ssp_cmat: Optional[np.ndarray] = None     # 'Q' [npts, nr]
    fg: tuple = (20.0, 35.0, 8.0, 0.0)        # Francois-Garrison T, S, pH, z_bar
    top: Halfspace = field(default_factory=lambda: Halfspace("V"))
    bot: Halfspace = field(default_factory=lambda: Halfspace("R"))
    ati: Boundary = field(default_factory=Boundary)
    bty: Boundary = field(default_factory=Boundary)
    trc: ReflTable = field(default_factory=ReflTable)
    brc: ReflTable = field(default_factory=ReflTable)
    Sz: np.ndarray = field(default_factory=lambda: np.zeros(1, np.float32))
    Rz: np.ndarray = field(default_factory=lambda: np.zeros(1, np.float32))
    Rr_km: np.ndarray = field(default_factory=lambda: np.zeros(1))
    alpha_deg: np.ndarray = field(default_factory=lambda: np.zeros(1))
    iSingle_alpha: int = 0
    run_type: str = "CG"                      # Beam->RunType (up to 7 chars)
    beam_type: str = "   "                    # Beam->Type[1..3]
    deltas: float = 0.0
    box_r_km: float = 101.0
    box_z: float = 5500.0
    epsMultiplier: float = 1.0
    rLoop: float = 1.0
    iBeamWindow: int = 4
    sbp: Optional[np.ndarray] = None          # [n, 2] deg / dB
    bio: Optional[np.ndarray] = None          # [n, 5] z1, z2 (m), f0 (Hz), Q, a0 (dB/km); atten_unit[1] == 'B'

    # ------------------------------------------------------------------
    @property
    def nrays(self):
        na = 1 if self.iSingle_alpha > 0 else len(self.alpha_deg)
        return na * len(self.Sz)

    @property
    def field_shape(self):
        nrz = 1 if (len(self.run_type) > 4 and self.run_type[4] == "I") else len(self.Rz)
        return (len(self.Sz), nrz, len(self.Rr_km))

    def to_c(self):
        """Return (Env2D_C, keepalive) -- the struct points into numpy arrays."""
        keep = []

        def dp(a):
            if a is None:
                return C.cast(None, _abi.c_double_p)
            arr = np.ascontiguousarray(a, dtype=np.float64)
            keep.append(arr)
            return arr.ctypes.data_as(_abi.c_double_p)

        def fp(a):
            arr = np.ascontiguousarray(a, dtype=np.float32)
            keep.append(arr)
            return arr.ctypes.data_as(_abi.c_float_p)

        e = _abi.Env2D_C()
        e.precision = self.precision
        e.title = self.title.encode()[:79]
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
        if self.ssp_type == "Q":
            e.ssp_nr = len(self.ssp_r_km)
            e.ssp_r_km = dp(self.ssp_r_km)
            e.ssp_cmat = dp(np.asarray(self.ssp_cmat, np.float64).reshape(-1))
        e.fg_T, e.fg_S, e.fg_pH, e.fg_z_bar = self.fg
        for dst, hs in ((e.top, self.top), (e.bot, self.bot)):
            dst.bc = hs.bc.encode()
            dst.depth, dst.alphaR, dst.betaR, dst.rho = hs.depth, hs.alphaR, hs.betaR, hs.rho
            dst.alphaI, dst.betaI, dst.zTemp, dst.Mz = hs.alphaI, hs.betaI, hs.zTemp, hs.Mz
        for dst, b in ((e.ati, self.ati), (e.bty, self.bty)):
            dst.type = (b.type + "  ")[:2].encode()
            dst.npts = len(b.r_km)
            if dst.npts:
                dst.r_km = dp(b.r_km)
                dst.z = dp(b.z)
                if b.type[1:2] == "L":
                    dst.alphaR, dst.betaR, dst.rho = dp(b.alphaR), dp(b.betaR), dp(b.rho)
                    dst.alphaI, dst.betaI = dp(b.alphaI), dp(b.betaI)
        for dst, t in ((e.trc, self.trc), (e.brc, self.brc)):
            dst.npts = len(t.theta)
            if dst.npts:
                dst.theta, dst.r, dst.phi = dp(t.theta), dp(t.r), dp(t.phi)
        e.NSz, e.NRz, e.NRr = len(self.Sz), len(self.Rz), len(self.Rr_km)
        e.Sz, e.Rz, e.Rr_km = fp(self.Sz), fp(self.Rz), dp(self.Rr_km)
        e.Nalpha = len(self.alpha_deg)
        e.iSingle_alpha = self.iSingle_alpha
        e.alpha_deg = dp(self.alpha_deg)
        e.RunType = (self.run_type + " " * 7)[:7].encode()
        e.BeamType = (self.beam_type + "   ")[:3].encode()
        e.deltas, e.box_r_km, e.box_z = self.deltas, self.box_r_km, self.box_z
        e.epsMultiplier, e.rLoop = self.epsMultiplier, self.rLoop
        e.iBeamWindow, e.Nimage = self.iBeamWindow, 1
        if self.sbp is not None:
            e.NSBPPts = len(self.sbp)
            e.sbp = dp(np.asarray(self.sbp, np.float64).reshape(-1))
        if self.bio is not None:
            e.NBioLayers = len(self.bio)
            e.bio = dp(np.asarray(self.bio, np.float64).reshape(-1))
        return e, keep


# ----------------------------------------------------------------------------
# Workloads
# ----------------------------------------------------------------------------

def munk_c(z):
    eta = 2.0 * (np.asarray(z, np.float64) - 1300.0) / 1300.0
    return 1500.0 * (1.0 + 0.00737 * (eta - 1.0 + np.exp(-eta)))


def munk_env(nbeams=2001, nrz=201, nrr=1001, precision=64, run_type="CG", ssp_type="C",
             freq=50.0, sd=1000.0, r_max_km=100.0, alpha=(-20.0, 20.0)):
    """Plain 2D Munk case: 51 c-linear knots 0-5000 m, vacuum surface, flat
    acousto-elastic halfspace bottom (cp 1600 m/s, rho 1.8, 0.8 dB/lambda),
    source 1000 m, 50 Hz, hat Cartesian beams over -20..20 deg, auto step,
    box 5500 m x 101 km, receivers nrz depths (0-5000 m) x nrr ranges (0-100 km)."""
    z = np.linspace(0.0, 5000.0, 51)
    return Env2D(
        title="Munk profile", precision=precision, freq0=freq, ssp_type=ssp_type,
        atten_unit="W ", ssp_z=z, ssp_c=munk_c(z),
        top=Halfspace("V", depth=0.0),
        bot=Halfspace("A", depth=5000.0, alphaR=1600.0, betaR=0.0, rho=1.8, alphaI=0.8,
                      zTemp=5000.0),
        Sz=np.array([sd], np.float32),
        Rz=subtab(0.0, 5000.0, nrz, np.float32),
        Rr_km=subtab(0.0, r_max_km, nrr),
        alpha_deg=subtab(alpha[0], alpha[1], nbeams),
        run_type=run_type, deltas=0.0, box_r_km=r_max_km + 1.0, box_z=5500.0)


def _smooth_noise(rng, n, scale, ncomp=6):
    x = np.linspace(0.0, 1.0, n)
    out = np.zeros(n)
    for k in range(1, ncomp + 1):
        out += rng.normal() / k * np.sin(np.pi * k * x + rng.uniform(0, 2 * np.pi))
    out /= max(np.max(np.abs(out)), 1e-12)
    return scale * out


def c2_env(seed=20260002, nbeams=20001, nrz=501, nrr=5001, precision=64, run_type="Cb",
           r_max_km=200.0):
    """2D range-dependent case: 40-knot spline SSP (linear 1480->1540 m/s plus a
    seeded smooth +-15 m/s perturbation), 500-segment bathymetry (3000 m base,
    2000 m Gaussian seamount sigma 10 km, 20 m rms seeded roughness), halfspace
    bottom with cp ~ U(1550, 1800), source 50 m, 100 Hz, ray-centred Gaussian
    beams over +-60 deg, receivers nrz x nrr to 200 km."""
    rng = np.random.default_rng(seed)
    zmax = 3100.0
    z = np.linspace(0.0, zmax, 40)
    c = 1480.0 + 60.0 * z / zmax + _smooth_noise(rng, 40, 15.0)
    nseg = 500
    r = np.linspace(0.0, r_max_km, nseg + 1)
    depth = 3000.0 - 2000.0 * np.exp(-0.5 * ((r - r_max_km / 2) / 10.0) ** 2)
    depth = depth + rng.normal(0.0, 20.0, r.shape)
    cp = float(rng.uniform(1550.0, 1800.0))
    return Env2D(
        title="range-dependent seamount", precision=precision, freq0=100.0, ssp_type="S",
        atten_unit="W ", ssp_z=z, ssp_c=c,
        top=Halfspace("V", depth=0.0),
        bot=Halfspace("A", depth=zmax, alphaR=cp, betaR=0.0, rho=1.8, alphaI=0.5, zTemp=zmax),
        bty=Boundary("LS", r_km=r, z=depth),
        Sz=np.array([50.0], np.float32),
        Rz=subtab(0.0, 3000.0, nrz, np.float32),
        Rr_km=subtab(0.0, r_max_km, nrr),
        alpha_deg=subtab(-60.0, 60.0, nbeams),
        run_type=run_type, deltas=0.0, box_r_km=r_max_km + 1.0, box_z=zmax + 100.0)


def sweep_envs(n=4096, seed=20260004, precision=64, scale=1.0):
    """Heterogeneous batched sweep: per problem the seed draws SSP knots 10-200
    (c-linear or PCHIP), flat or 1-500-segment bathymetry, water depth
    U(100, 5000) m, frequency log-U(25, 1000) Hz, beams 500-20000, receiver grid
    51-501 depths x 201-5001 ranges. `scale` < 1 shrinks beams and grids
    proportionally (for CPU-sized tests)."""
    rng = np.random.default_rng(seed)
    envs = []
    for _ in range(n):
        nk = int(rng.integers(10, 201))
        sspt = "C" if rng.random() < 0.5 else "P"
        D = float(rng.uniform(100.0, 5000.0))
        z = np.linspace(0.0, D, nk)
        c0 = munk_c(z * (5000.0 / D))
        c = c0 + _smooth_noise(rng, nk, 5.0)
        flat = rng.random() < 0.3
        rmax = float(rng.uniform(5.0, 50.0))
        if flat:
            bty = Boundary()
        else:
            ns = int(rng.integers(1, 501))
            rb = np.linspace(0.0, rmax, ns + 1)
            db = D * (0.8 + 0.15 * np.sin(2 * np.pi * rb / rmax * rng.uniform(0.5, 3.0)))
            db = np.minimum(db, D * 0.999)
            bty = Boundary("LS", r_km=rb, z=db)
        freq = float(math.exp(rng.uniform(math.log(25.0), math.log(1000.0))))
        nb = max(8, int(rng.integers(500, 20001) * scale))
        nrz = max(3, int(rng.integers(51, 502) * scale))
        nrr = max(3, int(rng.integers(201, 5002) * scale))
        sd = float(rng.uniform(0.05, 0.6) * D)
        envs.append(Env2D(
            title="sweep", precision=precision, freq0=freq, ssp_type=sspt, atten_unit="W ",
            ssp_z=z, ssp_c=c, top=Halfspace("V", depth=0.0),
            bot=Halfspace("A", depth=D, alphaR=float(rng.uniform(1550, 1900)), betaR=0.0,
                          rho=1.8, alphaI=0.5, zTemp=D),
            bty=bty, Sz=np.array([sd], np.float32), Rz=subtab(0.0, D, nrz, np.float32),
            Rr_km=subtab(0.0, rmax, nrr), alpha_deg=subtab(-30.0, 30.0, nb),
            run_type="CG", deltas=0.0, box_r_km=rmax + 1.0, box_z=D + 100.0))
    return envs
