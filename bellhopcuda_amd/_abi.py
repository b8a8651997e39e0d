"""ctypes mirror of include/bhc_hip.h (the C ABI of the HIP tracer library).

Field order and types follow the header exactly; tests/test_abi.py checks the
sizes of the record types against the reference layouts (Arrival 40 B,
EigenHit 36 B, rayPt<false> 112 B).
"""
import ctypes as C

c_double_p = C.POINTER(C.c_double)
c_float_p = C.POINTER(C.c_float)
c_int32_p = C.POINTER(C.c_int32)


class Arrival(C.Structure):
    _fields_ = [("NTopBnc", C.c_int32), ("NBotBnc", C.c_int32),
                ("SrcDeclAngle", C.c_float), ("SrcAzimAngle", C.c_float),
                ("RcvrDeclAngle", C.c_float), ("RcvrAzimAngle", C.c_float),
                ("a", C.c_float), ("Phase", C.c_float), ("delay", C.c_float * 2)]


class EigenHit(C.Structure):
    _fields_ = [(n, C.c_int32) for n in
                ("is_", "itheta", "ir", "iz", "isx", "isy", "isz", "ialpha", "ibeta")]


class RayPt2D(C.Structure):
    _fields_ = [("p", C.c_double * 2), ("q", C.c_double * 2),
                ("NumTopBnc", C.c_int32), ("NumBotBnc", C.c_int32),
                ("x", C.c_double * 2), ("t", C.c_double * 2),
                ("c", C.c_double), ("Amp", C.c_double), ("Phase", C.c_double),
                ("tau", C.c_double * 2)]


class HalfspaceIn(C.Structure):
    _fields_ = [("bc", C.c_char), ("pad", C.c_char * 7), ("depth", C.c_double),
                ("alphaR", C.c_double), ("betaR", C.c_double), ("rho", C.c_double),
                ("alphaI", C.c_double), ("betaI", C.c_double), ("zTemp", C.c_double),
                ("Mz", C.c_double)]


class BdryIn(C.Structure):
    _fields_ = [("type", C.c_char * 2), ("pad", C.c_char * 6), ("npts", C.c_int32),
                ("pad2", C.c_int32), ("r_km", c_double_p), ("z", c_double_p),
                ("alphaR", c_double_p), ("betaR", c_double_p), ("rho", c_double_p),
                ("alphaI", c_double_p), ("betaI", c_double_p)]


class ReflIn(C.Structure):
    _fields_ = [("npts", C.c_int32), ("pad", C.c_int32), ("theta", c_double_p),
                ("r", c_double_p), ("phi", c_double_p)]


class Env2D_C(C.Structure):
    _fields_ = [
        ("precision", C.c_int32), ("pad0", C.c_int32), ("title", C.c_char * 80),
        ("freq0", C.c_double),
        ("ssp_type", C.c_char), ("atten_unit", C.c_char * 2), ("pad1", C.c_char * 5),
        ("ssp_npts", C.c_int32), ("ssp_nr", C.c_int32),
        ("ssp_z", c_double_p), ("ssp_alphaR", c_double_p), ("ssp_alphaI", c_double_p),
        ("ssp_rho", c_double_p), ("ssp_betaR", c_double_p), ("ssp_betaI", c_double_p),
        ("ssp_r_km", c_double_p), ("ssp_cmat", c_double_p),
        ("fg_T", C.c_double), ("fg_S", C.c_double), ("fg_pH", C.c_double),
        ("fg_z_bar", C.c_double),
        ("top", HalfspaceIn), ("bot", HalfspaceIn),
        ("ati", BdryIn), ("bty", BdryIn), ("trc", ReflIn), ("brc", ReflIn),
        ("NSz", C.c_int32), ("NRz", C.c_int32), ("NRr", C.c_int32), ("pad2", C.c_int32),
        ("Sz", c_float_p), ("Rz", c_float_p), ("Rr_km", c_double_p),
        ("Nalpha", C.c_int32), ("iSingle_alpha", C.c_int32), ("alpha_deg", c_double_p),
        ("RunType", C.c_char * 7), ("BeamType", C.c_char * 3), ("pad3", C.c_char * 6),
        ("deltas", C.c_double), ("box_r_km", C.c_double), ("box_z", C.c_double),
        ("epsMultiplier", C.c_double), ("rLoop", C.c_double),
        ("iBeamWindow", C.c_int32), ("Nimage", C.c_int32),
        ("NSBPPts", C.c_int32), ("pad4", C.c_int32), ("sbp", c_double_p),
        ("NBioLayers", C.c_int32), ("pad5", C.c_int32), ("bio", c_double_p),
    ]


class BdryNode2D(C.Structure):
    _fields_ = [("x", C.c_double * 2), ("t", C.c_double * 2), ("n", C.c_double * 2),
                ("Noden", C.c_double * 2), ("Nodet", C.c_double * 2),
                ("Len", C.c_double), ("kappa", C.c_double),
                ("cP", C.c_double * 2), ("cS", C.c_double * 2), ("rho", C.c_double)]


class HS2D(C.Structure):
    _fields_ = [("bc", C.c_char), ("pad", C.c_char * 7), ("cP", C.c_double * 2),
                ("cS", C.c_double * 2), ("rho", C.c_double), ("depth", C.c_double)]


class Problem2D(C.Structure):
    _fields_ = [
        ("precision", C.c_int32), ("RunType", C.c_char * 7), ("BeamType", C.c_char * 4),
        ("ssp_type", C.c_char), ("pad0", C.c_int32), ("freq0", C.c_double),
        ("ssp_npts", C.c_int32), ("ssp_nr", C.c_int32),
        ("ssp_z", c_double_p), ("ssp_c", c_double_p), ("ssp_cz", c_double_p),
        ("ssp_n2", c_double_p), ("ssp_n2z", c_double_p), ("ssp_coef", c_double_p),
        ("ssp_rho", c_double_p), ("ssp_seg_r", c_double_p), ("ssp_cmat", c_double_p),
        ("ssp_czmat", c_double_p),
        ("top_hs", HS2D), ("bot_hs", HS2D),
        ("ati_type", C.c_char * 2), ("bty_type", C.c_char * 2),
        ("ati_npts", C.c_int32), ("bty_npts", C.c_int32),
        ("ati", C.POINTER(BdryNode2D)), ("bty", C.POINTER(BdryNode2D)),
        ("trc_npts", C.c_int32), ("brc_npts", C.c_int32),
        ("trc", c_double_p), ("brc", c_double_p),
        ("NSz", C.c_int32), ("NRz", C.c_int32), ("NRr", C.c_int32),
        ("NRz_per_range", C.c_int32),
        ("Sz", c_float_p), ("Rz", c_float_p), ("Rr", c_double_p), ("Delta_r", C.c_double),
        ("Nalpha", C.c_int32), ("iSingle_alpha", C.c_int32),
        ("alpha", c_double_p), ("alpha_cos", c_double_p), ("alpha_sin", c_double_p),
        ("Dalpha", C.c_double),
        ("deltas", C.c_double), ("box_r", C.c_double), ("box_z", C.c_double),
        ("epsMultiplier", C.c_double), ("rLoop", C.c_double),
        ("iBeamWindow", C.c_int32), ("Nimage", C.c_int32),
        ("NSBPPts", C.c_int32), ("pad1", C.c_int32), ("sbp", c_double_p),
    ]


class FieldOut(C.Structure):
    _fields_ = [
        ("field", c_float_p),
        ("arr", C.POINTER(Arrival)), ("narr", c_int32_p),
        ("max_narr", C.c_int32), ("allow_merging", C.c_int32),
        ("hits", C.POINTER(EigenHit)), ("neigen", c_int32_p),
        ("eigen_memsize", C.c_int32), ("pad", C.c_int32),
        ("ray_nsteps", c_int32_p), ("ray_ntop", c_int32_p), ("ray_nbot", c_int32_p),
        ("ray_xend", c_double_p),
        ("err", C.c_uint32), ("warn", C.c_uint32), ("err_count", C.c_uint32),
        ("warn_count", C.c_uint32),
        ("total_steps", C.c_int64), ("total_contrib", C.c_int64),
    ]


class RayOut(C.Structure):
    _fields_ = [("points", C.POINTER(RayPt2D)), ("nsteps", c_int32_p),
                ("max_points_per_ray", C.c_int32), ("pad", C.c_int32),
                ("err", C.c_uint32), ("warn", C.c_uint32), ("err_count", C.c_uint32),
                ("warn_count", C.c_uint32)]


class Bdry3DIn(C.Structure):
    _fields_ = [("type", C.c_char * 2), ("pad", C.c_char * 6), ("nx", C.c_int32), ("ny", C.c_int32),
                ("x_km", c_double_p), ("y_km", c_double_p), ("z", c_double_p),
                ("province", c_int32_p), ("nprov", C.c_int32), ("pad2", C.c_int32), ("prov", c_double_p)]


class Env3D_C(C.Structure):
    _fields_ = [
        ("precision", C.c_int32), ("pad0", C.c_int32), ("freq0", C.c_double),
        ("ssp_type", C.c_char), ("atten_unit", C.c_char * 2), ("pad1", C.c_char * 5),
        ("ssp_npts", C.c_int32), ("pad2", C.c_int32),
        ("ssp_z", c_double_p), ("ssp_alphaR", c_double_p), ("ssp_alphaI", c_double_p),
        ("ssp_rho", c_double_p), ("ssp_betaR", c_double_p), ("ssp_betaI", c_double_p),
        ("hex_nx", C.c_int32), ("hex_ny", C.c_int32), ("hex_nz", C.c_int32), ("pad3", C.c_int32),
        ("hex_x_km", c_double_p), ("hex_y_km", c_double_p), ("hex_z", c_double_p), ("hex_c", c_double_p),
        ("fg_T", C.c_double), ("fg_S", C.c_double), ("fg_pH", C.c_double), ("fg_z_bar", C.c_double),
        ("top", HalfspaceIn), ("bot", HalfspaceIn),
        ("ati", Bdry3DIn), ("bty", Bdry3DIn), ("trc", ReflIn), ("brc", ReflIn),
        ("NSx", C.c_int32), ("NSy", C.c_int32), ("NSz", C.c_int32), ("NRz", C.c_int32),
        ("NRr", C.c_int32), ("Ntheta", C.c_int32),
        ("Sx_km", c_double_p), ("Sy_km", c_double_p), ("Sz", c_float_p), ("Rz", c_float_p),
        ("Rr_km", c_double_p), ("theta_deg", c_double_p),
        ("Nalpha", C.c_int32), ("iSingle_alpha", C.c_int32), ("Nbeta", C.c_int32),
        ("iSingle_beta", C.c_int32), ("alpha_deg", c_double_p), ("beta_deg", c_double_p),
        ("RunType", C.c_char * 7), ("BeamType", C.c_char * 3), ("pad4", C.c_char * 6),
        ("deltas", C.c_double), ("box_x_km", C.c_double), ("box_y_km", C.c_double), ("box_z", C.c_double),
        ("NSBPPts", C.c_int32), ("pad5", C.c_int32), ("sbp", c_double_p),
        ("NBioLayers", C.c_int32), ("pad6", C.c_int32), ("bio", c_double_p),
    ]


class BdryNode3D(C.Structure):
    _fields_ = [("x", C.c_double * 3), ("n1", C.c_double * 3), ("n2", C.c_double * 3), ("Noden", C.c_double * 3),
                ("z_xx", C.c_double), ("z_xy", C.c_double), ("z_yy", C.c_double),
                ("kappa_xx", C.c_double), ("kappa_xy", C.c_double), ("kappa_yy", C.c_double),
                ("province", C.c_int32), ("pad", C.c_int32)]


class Problem3D(C.Structure):
    _fields_ = [
        ("precision", C.c_int32), ("RunType", C.c_char * 7), ("BeamType", C.c_char * 4),
        ("ssp_type", C.c_char), ("pad0", C.c_int32), ("freq0", C.c_double),
        ("ssp_npts", C.c_int32), ("pad1", C.c_int32),
        ("ssp_z", c_double_p), ("ssp_c", c_double_p), ("ssp_cz", c_double_p), ("ssp_n2", c_double_p),
        ("ssp_n2z", c_double_p), ("ssp_coef", c_double_p), ("ssp_rho", c_double_p),
        ("hex_nx", C.c_int32), ("hex_ny", C.c_int32), ("hex_nz", C.c_int32), ("pad2", C.c_int32),
        ("hex_x", c_double_p), ("hex_y", c_double_p), ("hex_z", c_double_p), ("hex_c", c_double_p),
        ("hex_cz", c_double_p),
        ("top_hs", HS2D), ("bot_hs", HS2D),
        ("ati_nx", C.c_int32), ("ati_ny", C.c_int32), ("bty_nx", C.c_int32), ("bty_ny", C.c_int32),
        ("ati", C.POINTER(BdryNode3D)), ("bty", C.POINTER(BdryNode3D)),
        ("ati_type", C.c_char * 2), ("bty_type", C.c_char * 2), ("ati_nprov", C.c_int32), ("bty_nprov", C.c_int32),
        ("ati_prov", c_double_p), ("bty_prov", c_double_p),
        ("trc_npts", C.c_int32), ("brc_npts", C.c_int32), ("trc", c_double_p), ("brc", c_double_p),
        ("NSx", C.c_int32), ("NSy", C.c_int32), ("NSz", C.c_int32), ("NRz", C.c_int32),
        ("NRr", C.c_int32), ("Ntheta", C.c_int32),
        ("Sx", c_double_p), ("Sy", c_double_p), ("Sz", c_float_p), ("Rz", c_float_p),
        ("Rr", c_double_p), ("t_rcvr", c_double_p), ("Delta_r", C.c_double),
        ("Nalpha", C.c_int32), ("iSingle_alpha", C.c_int32), ("Nbeta", C.c_int32),
        ("iSingle_beta", C.c_int32),
        ("alpha", c_double_p), ("alpha_cos", c_double_p), ("alpha_sin", c_double_p),
        ("beta", c_double_p), ("beta_cos", c_double_p), ("beta_sin", c_double_p),
        ("Dalpha", C.c_double), ("Dbeta", C.c_double),
        ("deltas", C.c_double), ("box_x", C.c_double), ("box_y", C.c_double), ("box_z", C.c_double),
        ("NSBPPts", C.c_int32), ("pad3", C.c_int32), ("sbp", c_double_p),
        ("theta", c_double_p),
    ]


class RayPt3D(C.Structure):
    _fields_ = [("p", C.c_double * 4), ("q", C.c_double * 4), ("phi", C.c_double), ("dummy", C.c_double),
                ("NumTopBnc", C.c_int32), ("NumBotBnc", C.c_int32),
                ("x", C.c_double * 3), ("t", C.c_double * 3),
                ("c", C.c_double), ("Amp", C.c_double), ("Phase", C.c_double), ("tau", C.c_double * 2)]


class RayOut3D(C.Structure):
    _fields_ = [("points", C.POINTER(RayPt3D)), ("nsteps", c_int32_p),
                ("max_points_per_ray", C.c_int32), ("pad", C.c_int32),
                ("err", C.c_uint32), ("warn", C.c_uint32), ("err_count", C.c_uint32),
                ("warn_count", C.c_uint32)]
