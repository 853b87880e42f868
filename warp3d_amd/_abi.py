"""ctypes mirror of include/w3d.h (structures, enums and prototypes).

The C header is the boundary; this module only restates its layouts so the
Python host code and the tests can fill a ``w3d_model`` from numpy arrays.
"""
import ctypes as C

import numpy as np

W3D_ABI_VERSION = 5
W3D_MAX_CURVES = 8

W3D_OK = 0
W3D_CUT_STEP = 1

W3D_Q3DISOP = 1
W3D_L3DISOP = 2
W3D_TET10 = 6

W3D_MM01 = 1
W3D_MM10 = 10

W3D_FP64 = 0
W3D_FP32 = 1

STRESS_N, STRESS_N1, STRAIN_N, STRAIN_N1, HIST_N, HIST_N1, CEP, ESTIFF, \
    EINFVEC, ELEM_VOL, ROT_N1 = range(11)

NNODE = {W3D_Q3DISOP: 20, W3D_L3DISOP: 8, W3D_TET10: 10}


class W3DMaterial(C.Structure):
    _fields_ = [
        ("model", C.c_int32),
        ("e", C.c_float), ("nu", C.c_float), ("beta", C.c_float),
        ("tan_e", C.c_float), ("yld_pt", C.c_float),
        ("linear_elastic", C.c_int32),
        ("slip_type", C.c_int32), ("elastic_type", C.c_int32),
        ("cp_e", C.c_double), ("cp_nu", C.c_double), ("cp_mu", C.c_double),
        ("harden_n", C.c_double), ("theta_0", C.c_double),
        ("tau_y", C.c_double), ("tau_v", C.c_double),
        ("voce_m", C.c_double), ("iD_v", C.c_double),
        ("eps_dot_0_y", C.c_double),
        ("atol", C.c_double), ("atol1", C.c_double),
        ("rtol", C.c_double), ("rtol1", C.c_double),
        ("miter", C.c_int32), ("alter_mode", C.c_int32),
        ("angle_type", C.c_int32), ("angle_convention", C.c_int32),
        ("ncrystals", C.c_int32), ("pad_", C.c_int32),
        ("alpha", C.c_float * 6), ("alpha_zero", C.c_float),
        ("curve_type", C.c_int32), ("ncurves", C.c_int32),
        ("curve_eps", C.c_double * 2),
        ("curve_temp", C.c_double * W3D_MAX_CURVES),
        ("curve_e", C.c_double * W3D_MAX_CURVES),
        ("curve_nu", C.c_double * W3D_MAX_CURVES),
        ("curve_alpha", C.c_double * W3D_MAX_CURVES),
        ("curve_sig", (C.c_double * 2) * W3D_MAX_CURVES),
    ]


def set_thermal(m, alpha, alpha_zero=0.0):
    """Thermal expansion of a material: alpha a scalar (isotropic: xx, yy,
    zz; shear terms zero) or the six props 9, 13, 34-37 (xx, yy, zz, xy,
    yz, xz), rounded through REAL like the reference's property table."""
    a = np.broadcast_to(np.asarray(alpha, float), (3,)) if np.ndim(alpha) == 0 \
        else np.asarray(alpha, float)
    vals = list(a) + [0.0] * (6 - len(a))
    for k in range(6):
        m.alpha[k] = vals[k]
    m.alpha_zero = alpha_zero
    return m


def set_curves(m, eps, curves):
    """A segmental curve set on an mm01 material: eps = (eps_1, eps_2), the
    plastic strains of the curve points; curves = [(sig_1, sig_2)] (one
    temperature-independent curve) or [(temp, e, nu, alpha, sig_1, sig_2)]
    with ascending temperatures (temperature dependent)."""
    m.curve_eps[0], m.curve_eps[1] = eps
    if len(curves[0]) == 2:
        m.curve_type = 1
        m.ncurves = 1
        m.curve_sig[0][0], m.curve_sig[0][1] = curves[0]
        return m
    if len(curves) > W3D_MAX_CURVES:
        raise ValueError("too many curves")
    m.curve_type = 2
    m.ncurves = len(curves)
    for k, (t, e, nu, al, s1, s2) in enumerate(curves):
        m.curve_temp[k] = t
        m.curve_e[k] = e
        m.curve_nu[k] = nu
        m.curve_alpha[k] = al
        m.curve_sig[k][0] = s1
        m.curve_sig[k][1] = s2
    return m


class W3DModel(C.Structure):
    _fields_ = [
        ("abi_version", C.c_int32),
        ("nonode", C.c_int32),
        ("noelem", C.c_int32),
        ("c", C.POINTER(C.c_double)),
        ("incid", C.POINTER(C.c_int32)),
        ("incmap", C.POINTER(C.c_int32)),
        ("etype", C.POINTER(C.c_int32)),
        ("int_order", C.POINTER(C.c_int32)),
        ("matnum", C.POINTER(C.c_int32)),
        ("bbar", C.POINTER(C.c_int32)),
        ("geonl", C.POINTER(C.c_int32)),
        ("angles", C.POINTER(C.c_double)),
        ("angle_stride", C.c_int32),
        ("nmat", C.c_int32),
        ("mats", C.POINTER(W3DMaterial)),
        ("nelblk", C.c_int32),
        ("elblks", C.POINTER(C.c_int32)),
        ("auto_size", C.c_int32),
        ("eps_bbar", C.c_double),
        ("trn", C.POINTER(C.c_int32)),
        ("trnmat", C.POINTER(C.c_double)),
    ]


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def lptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


def mm01_material(e, nu, yld_pt, tan_e=None, hprime=None, beta=1.0,
                  linear_elastic=False):
    """An mm01 material.  Give either the tangent modulus tan_e (the
    reference's input) or a target H' (then tan_e = E H / (E + H), rounded
    to REAL like the reference's input would be)."""
    m = W3DMaterial()
    m.model = W3D_MM01
    m.e = e
    m.nu = nu
    m.yld_pt = yld_pt
    if tan_e is None:
        tan_e = float(np.float32(e * hprime / (e + hprime)))
    m.tan_e = tan_e
    m.beta = beta
    m.linear_elastic = int(bool(linear_elastic))
    return m


def mm10_material(tau_y, tau_v, theta_0, voce_m=1.0, harden_n=20.0,
                  slip_type="fcc", elastic="isotropic", e=69000.0, nu=0.33,
                  mu=None, iD_v=0.0, ncrystals=1, angle_type="degrees",
                  atol=1.0e-5, atol1=1.0e-5, rtol=5.0e-5, rtol1=1.0e-5,
                  miter=30, alter_mode=False, eps_dot_0_y=1.0e10):
    """An mm10 (crystal plasticity, Voce hardening) material.  Defaults are
    the reference's crystal defaults (mod_crystals.f:238-417)."""
    m = W3DMaterial()
    m.model = W3D_MM10
    m.slip_type = {"fcc": 1, "bcc": 2}[slip_type]
    m.elastic_type = {"isotropic": 1, "cubic": 2}[elastic]
    m.cp_e = e
    m.cp_nu = nu
    m.cp_mu = e / 2.0 / (1.0 + nu) if mu is None else mu
    m.harden_n = harden_n
    m.theta_0 = theta_0
    m.tau_y = tau_y
    m.tau_v = tau_v
    m.voce_m = voce_m
    m.iD_v = iD_v
    m.eps_dot_0_y = eps_dot_0_y
    m.atol, m.atol1, m.rtol, m.rtol1 = atol, atol1, rtol, rtol1
    m.miter = miter
    m.alter_mode = int(bool(alter_mode))
    m.angle_type = {"degrees": 1, "radians": 2}[angle_type]
    m.angle_convention = 1
    m.ncrystals = ncrystals
    return m


def mm10_hist_size(ncrystals=1):
    """History words of one mm10 point: 87 common + 71 per crystal
    (mm10_d.f:25)."""
    return 87 + 71 * ncrystals


class ModelArrays:
    """Holds the numpy arrays a w3d_model points into (keeps them alive)."""

    def __init__(self, coords, incid, incmap, etype, int_order, matnum,
                 bbar, materials, elblks=None, auto_size=128, eps_bbar=0.0,
                 angles=None, angle_stride=1, geonl=None, trnmat=None):
        """trnmat: None, or a dict {node (1-based): 3x3 T} of skewed
        constraint systems (u_c = T u_g)."""
        self.coords = np.ascontiguousarray(coords, dtype=np.float64).ravel()
        self.incid = np.ascontiguousarray(incid, dtype=np.int32).ravel()
        self.incmap = np.ascontiguousarray(incmap, dtype=np.int32)
        self.etype = np.ascontiguousarray(etype, dtype=np.int32)
        self.int_order = np.ascontiguousarray(int_order, dtype=np.int32)
        self.matnum = np.ascontiguousarray(matnum, dtype=np.int32)
        self.bbar = np.ascontiguousarray(bbar, dtype=np.int32)
        self.geonl = (None if geonl is None else
                      np.ascontiguousarray(geonl, dtype=np.int32))
        self.angles = (None if angles is None else
                       np.ascontiguousarray(angles, dtype=np.float64).ravel())
        self.angle_stride = angle_stride
        self.materials = list(materials)
        self.mats = (W3DMaterial * len(self.materials))(*self.materials)
        self.elblks = (None if elblks is None else
                       np.ascontiguousarray(elblks, dtype=np.int32).ravel())
        self.auto_size = auto_size
        self.eps_bbar = eps_bbar
        self.nonode = self.coords.size // 3
        self.noelem = self.etype.size
        self.trn = self.trnmat = None
        if trnmat:
            self.trn = np.zeros(self.nonode, np.int32)
            self.trnmat = np.zeros(9 * self.nonode)
            for n, t in trnmat.items():
                self.trn[n - 1] = 1
                self.trnmat[9 * (n - 1):9 * n] = np.asarray(
                    t, float).reshape(3, 3).ravel(order="F")

    def struct(self):
        m = W3DModel()
        m.abi_version = W3D_ABI_VERSION
        m.nonode = self.nonode
        m.noelem = self.noelem
        m.c = dptr(self.coords)
        m.incid = iptr(self.incid)
        m.incmap = iptr(self.incmap)
        m.etype = iptr(self.etype)
        m.int_order = iptr(self.int_order)
        m.matnum = iptr(self.matnum)
        m.bbar = iptr(self.bbar)
        m.geonl = (iptr(self.geonl) if self.geonl is not None
                   else C.POINTER(C.c_int32)())
        m.angles = (dptr(self.angles) if self.angles is not None
                    else C.POINTER(C.c_double)())
        m.angle_stride = self.angle_stride
        m.nmat = len(self.materials)
        m.mats = C.cast(self.mats, C.POINTER(W3DMaterial))
        if self.elblks is not None:
            m.nelblk = self.elblks.size // 4
            m.elblks = iptr(self.elblks)
        else:
            m.nelblk = 0
            m.elblks = C.POINTER(C.c_int32)()
        m.auto_size = self.auto_size
        m.eps_bbar = self.eps_bbar
        if self.trn is not None:
            m.trn = iptr(self.trn)
            m.trnmat = dptr(self.trnmat)
        else:
            m.trn = C.POINTER(C.c_int32)()
            m.trnmat = C.POINTER(C.c_double)()
        return m


# prototypes shared by the device library and the oracle (same signatures
# with a different prefix)
_P = C.c_void_p
_I = C.c_int32
_D = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int32)
_LP = C.POINTER(C.c_int64)

class W3DNewtonOpts(C.Structure):
    """w3d_newton_opts: Newton and solver controls of a served problem."""
    _fields_ = [("nsteps", C.c_int32), ("max_iters", C.c_int32),
                ("precision", C.c_int32), ("pcg_maxit", C.c_int32),
                ("adaptive", C.c_int32), ("reserved", C.c_int32),
                ("tol", C.c_double), ("tol_abs", C.c_double),
                ("pcg_rtol", C.c_double)]


W3D_PROBLEM_RUNNING = 0
W3D_PROBLEM_STEP_DONE = 1
W3D_PROBLEM_DONE = 2
W3D_PROBLEM_FAILED = 3

PROTOS = {
    "get_blocking": (_I, [_P, _IP, _IP]),
    "block_info": (_I, [_P, _I, _IP]),
    "recover": (_I, [_P, _I, _I, _I, _D, _D, _D, _D, _D]),
    "recover_blocks": (_I, [_P, _I, _I, _I, _I, _I, _D, _D]),
    "tangent_blocks": (_I, [_P, _I, _I, _I, _I]),
    "pattern": (_I, [_P, _IP, _IP, _LP]),
    "get_pattern": (_I, [_P, _IP, _IP, _IP, _IP]),
    "assemble_vss": (_I, [_P, _D, _D]),
    "commit_step": (_I, [_P]),
    "rollback": (_I, [_P]),
    "get_block_state": (_I, [_P, _I, _I, _D, C.c_int64]),
    "set_block_state": (_I, [_P, _I, _I, _D, C.c_int64]),
    "set_temperatures": (_I, [_P, _D, _D, _D, _D, _I]),
}

# the library's own forms of the per-block calls (the oracle's
# orc_tangent_blocks takes no precision) and its per-block interface
LIB_BLOCK_PROTOS = {
    "tangent_blocks": (_I, [_P, _I, _I, _I, _I, _I]),
    "set_displacements": (_I, [_P, _D, _D]),
    "recover_block": (_I, [_P, _I, _I, _I, _I, _D, _D, _IP]),
    "assemble_ifv": (_I, [_P, _D, _D, _LP]),
    "get_ifv_sums": (_I, [_P, _D, _LP]),
    "tangent_block": (_I, [_P, _I, _I, _I, _I]),
}

# asymmetric assembly (the library; the oracle's counterparts take the CSR
# pattern from the caller)
ASYM_PROTOS = {
    "set_asymmetric": (_I, [_P, _I]),
    "get_csr_pattern": (_I, [_P, _LP, _LP, _IP]),
    "assemble_csr": (_I, [_P, _D]),
    "assemble_csr_dev": (_I, [_P, _P]),
    "get_block_ek_full": (_I, [_P, _I, _D, C.c_int64]),
    "solve_bicgstab": (_I, [_P, _D, _D, C.c_double, _I, _IP,
                            C.POINTER(C.c_double)]),
}

# the batched serving entry points (the library only; the oracle mirrors
# the block interface above)
SERVE_PROTOS = {
    "problem_create": (_I, [_P, _IP, _D, C.POINTER(W3DNewtonOpts),
                            C.POINTER(_P)]),
    "problem_destroy": (_I, [_P]),
    "problem_status": (_I, [_P, _IP, _IP, _IP]),
    "problem_step_record": (_I, [_P, _I, _IP, _IP, _IP, _D, _D, _D]),
    "problem_get_u": (_I, [_P, _D]),
    "problem_device_bytes": (_I, [_P, _LP]),
    "batch_create": (_I, [_I, C.POINTER(_P)]),
    "batch_destroy": (_I, [_P]),
    "batch_advance": (_I, [_P, C.POINTER(_P), _I, _IP]),
    "batch_phase_times": (_I, [_P, _D, _LP, _LP]),
}
