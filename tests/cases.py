"""Feature environments shared by the CPU core tests and the GPU parity tests:
each switches on one part of the path the Munk and seamount cases leave off."""
import numpy as np

from bellhopcuda_amd.env import Env2D, Halfspace, Boundary, ReflTable, subtab, munk_env, munk_c


def _base(**kw):
    return munk_env(nbeams=kw.pop("nbeams", 201), nrz=kw.pop("nrz", 41), nrr=kw.pop("nrr", 101), **kw)


def quad_ssp(precision=64):
    """Range-dependent quad SSP: the Munk profile at five ranges, perturbed;
    the profiles span the beam box so that no ray leaves the SSP."""
    e = _base(precision=precision, ssp_type="Q", r_max_km=40.0)
    e.ssp_r_km = np.array([-5.0, 0.0, 12.0, 25.0, 45.0])
    z = e.ssp_z
    e.ssp_cmat = np.stack([munk_c(z) + 8.0 * k * np.sin(np.pi * z / 5000.0) for k in range(5)], axis=1)
    e.box_r_km = 41.0
    return e


def curvilinear_bathymetry(precision=64):
    e = _base(precision=precision, r_max_km=60.0)
    r = np.linspace(0.0, 61.0, 41)
    e.bty = Boundary("CS", r_km=r, z=4800.0 - 1500.0 * np.exp(-0.5 * ((r - 30.0) / 8.0) ** 2))
    e.box_r_km = 61.0
    return e


def long_format_bathymetry(precision=64):
    e = _base(precision=precision, r_max_km=60.0)
    r = np.linspace(0.0, 61.0, 25)
    n = len(r)
    e.bty = Boundary("LL", r_km=r, z=4900.0 - 300.0 * np.sin(r / 9.0) ** 2,
                     alphaR=np.linspace(1550.0, 1750.0, n), betaR=np.zeros(n), rho=np.linspace(1.5, 2.0, n),
                     alphaI=np.full(n, 0.3), betaI=np.zeros(n))
    e.box_r_km = 61.0
    return e


def altimetry(precision=64):
    e = _base(precision=precision, r_max_km=50.0)
    r = np.linspace(0.0, 51.0, 101)
    e.ati = Boundary("LS", r_km=r, z=4.0 + 3.0 * np.sin(2 * np.pi * r / 7.0))
    e.box_r_km = 51.0
    return e


def reflection_table(precision=64):
    e = _base(precision=precision)
    th = np.array([0.0, 10.0, 20.0, 40.0, 60.0, 90.0])
    e.bot = Halfspace("F", depth=5000.0)
    e.brc = ReflTable(th, np.array([1.0, 0.95, 0.8, 0.5, 0.3, 0.2]), np.array([180.0, 170.0, 150.0, 120.0, 90.0, 60.0]))
    return e


def elastic_bottom(precision=64):
    e = _base(precision=precision)
    e.bot = Halfspace("A", depth=5000.0, alphaR=1800.0, betaR=400.0, rho=1.9, alphaI=0.4, betaI=1.0, zTemp=5000.0)
    return e


def rigid_bottom_gauss(precision=64):
    e = _base(precision=precision, run_type="CB")
    e.bot = Halfspace("R", depth=5000.0)
    return e


def several_sources(precision=64):
    e = _base(precision=precision)
    e.Sz = np.array([200.0, 1000.0, 3000.0], np.float32)
    return e


def irregular_grid(precision=64):
    e = _base(precision=precision, run_type="CG RI")
    e.Rz = subtab(100.0, 4900.0, 101, np.float32)
    e.Rr_km = subtab(1.0, 100.0, 101)
    return e


def line_source(precision=64):
    return _base(precision=precision, run_type="CG X")


def beam_pattern(precision=64):
    e = _base(precision=precision)
    e.sbp = np.array([[-90.0, -20.0], [-10.0, -3.0], [0.0, 0.0], [10.0, -3.0], [90.0, -20.0]])
    return e


def attenuation_units(precision=64, unit="F "):
    """Volume attenuation in another unit, at a level that leaves the far
    field above the TL floor."""
    e = _base(precision=precision)
    e.atten_unit = unit
    level = {"F": 0.002, "M": 2e-5, "m": 2e-5, "Q": 5000.0, "L": 1e-6, "N": 2e-6, "W": 0.01}[unit[0]]
    e.ssp_alphaI = np.full(len(e.ssp_z), level)
    return e


def biological_layers(precision=64):
    """Biological volume attenuation ('B', Diachok's resonant layers) on top
    of dB/wavelength attenuation: two layers, one with its resonance near the
    run frequency."""
    e = _base(precision=precision)
    e.atten_unit = "WB"
    e.ssp_alphaI = np.full(len(e.ssp_z), 0.01)
    e.bio = np.array([[0.0, 350.0, 48.0, 5.0, 0.02], [900.0, 1600.0, 200.0, 20.0, 0.01]])
    return e


def beam_shift(precision=64):
    """Beam displacement on reflection from a fluid halfspace (RunType[6] 'S'),
    with hat beams at the Munk fan's angles."""
    e = _base(precision=precision, run_type="CG    S")
    return e


def beam_shift_elastic(precision=64):
    e = elastic_bottom(precision)
    e.run_type = "CG    S"
    return e


FEATURES = {
    "quad_ssp": quad_ssp,
    "curvilinear_bathymetry": curvilinear_bathymetry,
    "long_format_bathymetry": long_format_bathymetry,
    "altimetry": altimetry,
    "reflection_table": reflection_table,
    "elastic_bottom": elastic_bottom,
    "rigid_bottom_gauss": rigid_bottom_gauss,
    "several_sources": several_sources,
    "irregular_grid": irregular_grid,
    "line_source": line_source,
    "beam_pattern": beam_pattern,
    "atten_F_thorp": lambda precision=64: attenuation_units(precision, "FT"),
    "atten_M": lambda precision=64: attenuation_units(precision, "M "),
    "atten_N": lambda precision=64: attenuation_units(precision, "N "),
    "atten_Q": lambda precision=64: attenuation_units(precision, "Q "),
    "atten_L_francois_garrison": lambda precision=64: attenuation_units(precision, "LF"),
    "atten_m_power_law": lambda precision=64: attenuation_units(precision, "m "),
    "atten_biological": biological_layers,
    "beam_shift": beam_shift,
    "beam_shift_elastic": beam_shift_elastic,
}
