"""Shared parity checks: GPU (or product core) results against the oracle.

Bars (from the task's parity definition):
  * per-ray step and bounce counts equal on >= 99.9 % of rays (FP64);
  * ray end points within 1e-9 relative (FP64);
  * TL within 0.1 dB where TL < 120 dB.
"""
import numpy as np


def tl_db(field):
    a = np.abs(field).astype(np.float64)
    with np.errstate(divide="ignore"):
        return -20.0 * np.log10(np.maximum(a, 1e-300))


def compare_field(o, g, scaled_o, scaled_g, fp64=True, min_exact=0.999, path_rtol=1e-9, tl_tol=0.1):
    rep = {}
    rep["steps_exact"] = float(np.mean(o.nsteps == g.nsteps))
    rep["bounces_exact"] = float(np.mean((o.ntop == g.ntop) & (o.nbot == g.nbot)))
    same = (o.nsteps == g.nsteps)
    scale = np.maximum(np.abs(o.xend), 1.0)
    rel = np.max(np.abs(o.xend - g.xend) / scale, axis=1)
    rep["xend_max_rel"] = float(np.max(rel[same])) if same.any() else 0.0
    tlo, tlg = tl_db(scaled_o), tl_db(scaled_g)
    m = tlo < 120.0
    rep["tl_max_db"] = float(np.max(np.abs(tlo - tlg)[m])) if m.any() else 0.0
    rep["tl_points"] = int(m.sum())
    ok = rep["steps_exact"] >= min_exact and rep["bounces_exact"] >= min_exact and rep["tl_max_db"] <= tl_tol
    if fp64:
        ok = ok and rep["xend_max_rel"] <= path_rtol
    return ok, rep


def env_context(**settings):
    """A device context created with the given BHC_* environment settings
    (read once, at context creation), e.g. env_context(BHC_SCATTER3_MAP=0)."""
    import os
    import bellhopcuda_amd as bh
    old = {k: os.environ.get(k) for k in settings}
    os.environ.update({k: str(v) for k, v in settings.items()})
    try:
        return bh.Context(0)
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


def small_segment_context(mb):
    """A device context whose segment buffer holds only `mb` MiB, so that a
    run is split into several trace / scatter passes, some of them retraced
    after a segment-buffer overflow."""
    return env_context(BHC_SEG_BUFFER_MB=mb)
