"""ctypes wrapper of the CPU oracle (oracle/libw3d_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker; the product path never uses it.
"""
import ctypes as C
import os

import numpy as np

from warp3d_amd._abi import (PROTOS, W3DMaterial, W3DModel, dptr, iptr,
                             lptr)

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libw3d_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run make -C oracle")
        L = C.CDLL(path)
        for name, (res, args) in PROTOS.items():
            f = getattr(L, "orc_" + name)
            f.restype = res
            f.argtypes = args
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(W3DModel)]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_last_error.restype = C.c_char_p
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_block_sums.argtypes = [C.c_void_p, C.c_int,
                                     C.POINTER(C.c_double)]
        L.orc_tangent.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_set_asymmetric.argtypes = [C.c_void_p, C.c_int]
        L.orc_assemble_csr.argtypes = [C.c_void_p, C.POINTER(C.c_int64),
                                       C.POINTER(C.c_int32),
                                       C.POINTER(C.c_double)]
        L.orc_get_ek_full.argtypes = [C.c_void_p, C.c_int,
                                      C.POINTER(C.c_double), C.c_int64]
        L.orc_ngp.restype = C.c_int
        L.orc_ngp.argtypes = [C.c_int, C.c_int]
        L.orc_getgpts.argtypes = [C.c_int, C.c_int, C.c_int] + \
            [C.POINTER(C.c_double)] * 4
        L.orc_derivs.argtypes = [C.c_int, C.c_double, C.c_double,
                                 C.c_double] + [C.POINTER(C.c_double)] * 3
        L.orc_jacob1.argtypes = [C.c_int] + [C.POINTER(C.c_double)] * 7
        L.orc_bmatrix.argtypes = [C.c_int] + [C.POINTER(C.c_double)] * 5 + \
            [C.c_double, C.POINTER(C.c_double)]
        L.orc_mm01_gp.restype = C.c_int
        L.orc_mm01_gp.argtypes = [C.c_int, C.c_double, C.c_double,
                                  C.c_double, C.c_double, C.c_int,
                                  C.c_double] + \
            [C.POINTER(C.c_double)] * 5 + [C.c_double, C.POINTER(C.c_double)]
        L.orc_shapef.argtypes = [C.c_int, C.c_double, C.c_double,
                                 C.c_double, C.POINTER(C.c_double)]
        L.orc_hprime.restype = C.c_float
        L.orc_hprime.argtypes = [C.c_float, C.c_float]
        _D = C.POINTER(C.c_double)
        L.orc_set_time_increment.argtypes = [C.c_void_p, C.c_double]
        L.orc_mm10_crystal.restype = C.c_int
        L.orc_mm10_crystal.argtypes = [C.POINTER(W3DMaterial), _D,
                                       C.POINTER(Crystal)]
        L.orc_mm10_gp.restype = C.c_int
        L.orc_mm10_gp.argtypes = [C.POINTER(W3DMaterial), C.POINTER(Crystal),
                                  C.c_int, C.c_int, C.c_double, _D, _D, _D,
                                  _D, _D, C.c_int, _D]
        L.orc_mm10_residual.argtypes = [C.POINTER(W3DMaterial),
                                        C.POINTER(Crystal), _D, _D,
                                        C.c_double, _D, _D, _D, _D]
        L.orc_mm10_calc_grads_geonl.restype = C.c_int
        L.orc_mm10_calc_grads_geonl.argtypes = [C.c_int, C.c_int, C.c_int,
                                                C.c_int, _D, _D, _D, _D]
        L.orc_mm10_rotation_matrix.argtypes = [_D, C.c_int, _D]
        L.orc_mm10_rt2rve.argtypes = [_D, _D]
        L.orc_mm10_rt2rvw.argtypes = [_D, _D]
        L.orc_mm10_expmw.argtypes = [_D, _D]
        L.orc_lsq4.argtypes = [C.c_int, _D, _D, _D]
        L.orc_tcomp1.argtypes = [C.c_int] + [_D] * 6
        L.orc_fcomp1.argtypes = [_D, _D, _D]
        L.orc_fcomp1.restype = C.c_int
        L.orc_evcmp1.argtypes = [_D, _D]
        L.orc_irscp1.argtypes = [_D, _D]
        L.orc_rtcmp1.argtypes = [_D, _D]
        L.orc_getrm1.argtypes = [_D, C.c_int, _D]
        L.orc_ctran1.argtypes = [_D, _D, _D, C.c_int, C.c_double]
        L.orc_kgstiff.argtypes = [C.c_int] + [_D] * 6 + \
            [C.c_double, C.c_double, _D]
        L.orc_mm10_calc_grads.argtypes = [C.c_int] * 5 + [_D, _D]
        _LIB = L
    return _LIB


class OracleError(RuntimeError):
    pass


class Crystal(C.Structure):
    """orc_crystal: one crystal's constants in the reference frame."""
    _fields_ = [("C", C.c_double * 36), ("ms", C.c_double * 72),
                ("qs", C.c_double * 36), ("g", C.c_double * 9),
                ("angles", C.c_double * 3)]

    def stiffness(self):
        return np.array(self.C).reshape(6, 6, order="F")

    def schmid(self):
        return (np.array(self.ms).reshape(12, 6),
                np.array(self.qs).reshape(12, 3))

    def rotation(self):
        return np.array(self.g).reshape(3, 3, order="F")


def _chk(rc, what):
    if rc < 0:
        raise OracleError(f"{what}: {lib().orc_last_error().decode()}")
    return rc


# ---------------------------------------------------------------- leaves
def getgpts(etype, order, gpn):
    d = [C.c_double() for _ in range(4)]
    lib().orc_getgpts(etype, order, gpn, *[C.byref(x) for x in d])
    return tuple(x.value for x in d)


def ngp(etype, order):
    return lib().orc_ngp(etype, order)


def derivs(etype, xi, eta, zeta, nnode):
    a = [np.zeros(nnode) for _ in range(3)]
    lib().orc_derivs(etype, xi, eta, zeta, *[dptr(x) for x in a])
    return a


def shapef(etype, xi, eta, zeta, nnode):
    n = np.zeros(nnode)
    lib().orc_shapef(etype, xi, eta, zeta, dptr(n))
    return n


def jacob1(nxi, neta, nzeta, ce):
    nn = len(nxi)
    jac = np.zeros(9)
    gama = np.zeros(9)
    dj = C.c_double()
    lib().orc_jacob1(nn, dptr(np.ascontiguousarray(nxi)),
                     dptr(np.ascontiguousarray(neta)),
                     dptr(np.ascontiguousarray(nzeta)),
                     dptr(np.ascontiguousarray(ce, dtype=np.float64)),
                     dptr(jac), C.byref(dj), dptr(gama))
    return jac.reshape(3, 3), dj.value, gama.reshape(3, 3)


def bmatrix(gama, nxi, neta, nzeta, vol=None, eps_bbar=0.0):
    nn = len(nxi)
    b = np.zeros((6, 3 * nn))
    g = np.ascontiguousarray(gama, dtype=np.float64).ravel()
    v = (C.POINTER(C.c_double)() if vol is None
         else dptr(np.ascontiguousarray(vol, dtype=np.float64).ravel()))
    lib().orc_bmatrix(nn, dptr(g), dptr(np.ascontiguousarray(nxi)),
                      dptr(np.ascontiguousarray(neta)),
                      dptr(np.ascontiguousarray(nzeta)), v, eps_bbar,
                      dptr(b))
    return b


def hprime(e, et):
    return lib().orc_hprime(e, et)


def mm01_gp(step, e, nu, beta, hp, yld, cgn, deps, hist, hist1=None,
            first_hist11=None, lnelas=False):
    cgn = np.array(cgn, dtype=np.float64)
    cgn1 = np.zeros(9)
    hist = np.array(hist, dtype=np.float64)
    hist1 = np.zeros(11) if hist1 is None else np.array(hist1, float)
    cep = np.zeros(36)
    f11 = hist[10] if first_hist11 is None else first_hist11
    rc = lib().orc_mm01_gp(step, e, nu, beta, hp, int(lnelas), yld,
                           dptr(cgn), dptr(cgn1),
                           dptr(np.ascontiguousarray(deps, dtype=float)),
                           dptr(hist), dptr(hist1), f11, dptr(cep))
    if rc != 0:
        raise OracleError("mm01 negative plastic multiplier")
    return cgn1, hist1, cep.reshape(6, 6)


# ---------------------------------------------------------------- mm10
def mm10_crystal(mat, angles):
    cr = Crystal()
    a = np.ascontiguousarray(angles, dtype=np.float64)
    if lib().orc_mm10_crystal(C.byref(mat), dptr(a), C.byref(cr)) != 0:
        raise OracleError("mm10 crystal set-up")
    return cr


def mm10_gp(mat, crystals, step, iter0_off, dt, uddt, urcs_n, hist_n,
            asym=False, rn1=None):
    """One mm10 point: returns (rc, urcs_n1, hist_n1, hist_n as used).
    rn1: [R n+1] as a 3x3 array (finite strains), None for identity."""
    ncry = len(crystals)
    arr = (Crystal * ncry)(*crystals)
    hn = np.array(hist_n, dtype=np.float64)
    hn1 = np.zeros(87 + 71 * ncry)
    u1 = np.zeros(9)
    rcm = (None if rn1 is None else
           np.ascontiguousarray(np.asarray(rn1, float).ravel(order="F")))
    rc = lib().orc_mm10_gp(C.byref(mat), C.cast(arr, C.POINTER(Crystal)),
                           step, int(iter0_off), dt,
                           dptr(np.ascontiguousarray(uddt, dtype=float)),
                           dptr(np.ascontiguousarray(urcs_n, dtype=float)),
                           dptr(u1), dptr(hn), dptr(hn1), int(asym),
                           None if rcm is None else dptr(rcm))
    return rc, u1, hn1, hn


def mm10_residual(mat, crystal, hist_crystal, uddt, dt, x, rn1=None):
    """Residual (7) and Jacobian (7x7) of the local problem; rn1: [R n+1]
    as a 3x3 array, None for identity."""
    R = np.zeros(7)
    J = np.zeros(49)
    hc = np.ascontiguousarray(hist_crystal, float)
    ud = np.ascontiguousarray(uddt, float)
    xx = np.ascontiguousarray(x, float)
    rcm = (None if rn1 is None else
           np.ascontiguousarray(np.asarray(rn1, float).ravel(order="F")))
    lib().orc_mm10_residual(C.byref(mat), C.byref(crystal), dptr(hc),
                            dptr(ud), dt, dptr(xx), dptr(R), dptr(J),
                            None if rcm is None else dptr(rcm))
    return R, J.reshape(7, 7, order="F")


def mm10_calc_grads_geonl(etype, order, hel, rot, jac):
    """grad Fe words (27) of one element of a finite-strain block: hel
    the element's n-history (ngp, hist), rot its [R n+1] (ngp, 3, 3), jac
    its n = 0 Jacobian at the last point (3x3).  Returns (bad, words)."""
    hel = np.ascontiguousarray(hel, float)
    ngp, hist = hel.shape
    rot = np.ascontiguousarray(
        np.stack([np.asarray(r, float).ravel(order="F") for r in rot]))
    jac = np.ascontiguousarray(np.asarray(jac, float).ravel())
    out = np.zeros(27)
    bad = lib().orc_mm10_calc_grads_geonl(etype, order, ngp, hist,
                                          dptr(hel), dptr(rot), dptr(jac),
                                          dptr(out))
    return bad, out


def mm10_rotation_matrix(angles, angle_type=1):
    r = np.zeros(9)
    lib().orc_mm10_rotation_matrix(
        dptr(np.ascontiguousarray(angles, float)), angle_type, dptr(r))
    return r.reshape(3, 3, order="F")


def mm10_rt2rve(rt):
    rv = np.zeros(36)
    lib().orc_mm10_rt2rve(dptr(np.asfortranarray(rt, float).ravel(order="F")),
                          dptr(rv))
    return rv.reshape(6, 6, order="F")


def lsq4(a, b):
    """min |A x - b| for A (m x 4), as DGELS('N')."""
    a = np.asfortranarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(4)
    lib().orc_lsq4(a.shape[0], dptr(a.ravel(order="F")), dptr(b), dptr(x))
    return x


def mm10_calc_grads(etype, order, ngp_, span, hist, hloc):
    """grad Fe words (27) of a small-strain single-crystal block from its
    n-history hloc (hist, ngp, span) in Fortran order."""
    h = np.asfortranarray(hloc, dtype=np.float64)
    out = np.zeros(27)
    lib().orc_mm10_calc_grads(etype, order, ngp_, span, hist,
                              dptr(h.ravel(order="F")), dptr(out))
    return out


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def tcomp1(gama, nxi, neta, nzeta, ue):
    """Displacement gradient theta(9) at a point (gama row-major)."""
    th = np.zeros(9)
    g, a, b, c, u = (_f64(gama).ravel(), _f64(nxi), _f64(neta), _f64(nzeta),
                     _f64(ue))
    lib().orc_tcomp1(a.size, dptr(g), dptr(a), dptr(b), dptr(c), dptr(u),
                     dptr(th))
    return th


def fcomp1(theta):
    """(F row-major 3x3, det F, error flag)."""
    f = np.zeros(9)
    df = np.zeros(1)
    t = _f64(theta)
    err = lib().orc_fcomp1(dptr(t), dptr(f), dptr(df))
    return f.reshape(3, 3), df[0], err


def evcmp1(c6):
    ev = np.zeros(3)
    c = _f64(c6)
    lib().orc_evcmp1(dptr(c), dptr(ev))
    return ev


def irscp1(f):
    ui = np.zeros(6)
    ff = _f64(f).ravel()
    lib().orc_irscp1(dptr(ff), dptr(ui))
    return ui


def rtcmp1(f):
    r = np.zeros(9)
    ff = _f64(f).ravel()
    lib().orc_rtcmp1(dptr(ff), dptr(r))
    return r.reshape(3, 3)


def getrm1(r, opt):
    q = np.zeros(36)
    rr = _f64(r).ravel()
    lib().orc_getrm1(dptr(rr), opt, dptr(q))
    return q.reshape(6, 6)


def ctran1(d, q, cs, qbar, wdj):
    dd = _f64(d).ravel().copy()
    qq, c = _f64(q).ravel(), _f64(cs)
    lib().orc_ctran1(dptr(dd), dptr(qq), dptr(c), int(qbar), wdj)
    return dd.reshape(6, 6)


def kgstiff(gama, nxi, neta, nzeta, sig, dj, w, vol=None):
    """Geometric stiffness of one point, packed upper column-major."""
    nn = len(nxi)
    T = 3 * nn
    ek = np.zeros(T * (T + 1) // 2)
    g, a, b, c, s = (_f64(gama).ravel(), _f64(nxi), _f64(neta), _f64(nzeta),
                     _f64(sig))
    v = _f64(vol).ravel() if vol is not None else None
    lib().orc_kgstiff(nn, dptr(g), dptr(a), dptr(b), dptr(c),
                      dptr(v) if v is not None else C.POINTER(C.c_double)(),
                      dptr(s), dj, w,
                      dptr(ek))
    return ek


def mm10_expmw(wv):
    a = np.zeros(9)
    lib().orc_mm10_expmw(dptr(np.ascontiguousarray(wv, float)), dptr(a))
    return a.reshape(3, 3, order="F")


# ---------------------------------------------------------------- model
class OracleModel:
    """The oracle behind the same calls as warp3d_amd.BlockModel."""

    def __init__(self, arrays, threads=0):
        self._arrays = arrays
        self._m = arrays.struct()
        lib().orc_set_threads(threads)
        self.h = lib().orc_create(C.byref(self._m))
        if not self.h:
            raise OracleError(lib().orc_last_error().decode())
        self.nonode = arrays.nonode
        nb = C.c_int32()
        lib().orc_get_blocking(self.h, C.byref(nb), None)
        self.nelblk = nb.value
        self.elblks = np.zeros(4 * self.nelblk, np.int32)
        lib().orc_get_blocking(self.h, C.byref(nb), iptr(self.elblks))
        self.elblks = self.elblks.reshape(self.nelblk, 4)
        self.neqns = None
        self.ncoeff = None

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def set_time_increment(self, dt):
        lib().orc_set_time_increment(self.h, dt)

    def set_temperatures(self, temper_nodes=None, dtemp_nodes=None,
                         temper_elems=None, dtemp_elems=None,
                         incremental=True):
        a = [None if v is None else np.ascontiguousarray(v, float)
             for v in (temper_nodes, dtemp_nodes, temper_elems, dtemp_elems)]
        _chk(lib().orc_set_temperatures(
            self.h, *[dptr(v) if v is not None else None for v in a],
            int(bool(incremental))), "set_temperatures")

    def ifv_sums(self):
        """addifv's (sum_ifv, num_term_ifv) of the last recovery, summed
        block by block."""
        s, n = 0.0, 0
        for b in range(1, self.nelblk + 1):
            s = s + self.block_sums(b)[2]
            info = self.block_info(b)
            n += info["span"] * info["totdof"]
        return s, n

    def block_info(self, blk):
        info = np.zeros(5, np.int32)
        _chk(lib().orc_block_info(self.h, blk, iptr(info)), "block_info")
        return dict(zip(("span", "ngp", "hist", "totdof", "utsz"),
                        info.tolist()))

    def recover(self, step, iter_, u, du, extrapolated=False):
        ifv = np.zeros(3 * self.nonode)
        en = C.c_double()
        pw = C.c_double()
        rc = _chk(lib().orc_recover(
            self.h, step, iter_, int(extrapolated),
            dptr(np.ascontiguousarray(u, float)),
            dptr(np.ascontiguousarray(du, float)), dptr(ifv),
            C.cast(C.byref(en), C.POINTER(C.c_double)),
            C.cast(C.byref(pw), C.POINTER(C.c_double))), "recover")
        return ifv, en.value, pw.value, rc

    def recover_blocks(self, first, n, step, iter_, u, du,
                       extrapolated=False):
        return _chk(lib().orc_recover_blocks(
            self.h, first, n, step, iter_, int(extrapolated),
            dptr(np.ascontiguousarray(u, float)),
            dptr(np.ascontiguousarray(du, float))), "recover_blocks")

    def block_sums(self, blk):
        """(energy, plastic work, sum |element force|) of block blk's last
        recovery (do_nleps_block's block sums, addifv's sum_ifv terms)."""
        out = np.zeros(3)
        _chk(lib().orc_block_sums(self.h, blk, dptr(out)), "block_sums")
        return tuple(out)

    def tangent(self, step, iter_, precision=0):
        return _chk(lib().orc_tangent(self.h, step, iter_), "tangent")

    def tangent_blocks(self, first, n, step, iter_):
        return _chk(lib().orc_tangent_blocks(self.h, first, n, step, iter_),
                    "tangent_blocks")

    def pattern(self, cstmap):
        neq = C.c_int32()
        nco = C.c_int64()
        _chk(lib().orc_pattern(self.h, iptr(np.ascontiguousarray(
            cstmap, np.int32)), C.byref(neq), C.byref(nco)), "pattern")
        self.neqns, self.ncoeff = neq.value, nco.value
        return self.neqns, self.ncoeff

    def get_pattern(self):
        d = np.zeros(3 * self.nonode, np.int32)
        e = np.zeros(self.neqns, np.int32)
        p = np.zeros(self.neqns, np.int32)
        k = np.zeros(max(self.ncoeff, 1), np.int32)
        _chk(lib().orc_get_pattern(self.h, iptr(d), iptr(e), iptr(p),
                                   iptr(k)), "get_pattern")
        return d, e, p, k[:self.ncoeff]

    def assemble_vss(self):
        kd = np.zeros(self.neqns)
        kc = np.zeros(max(self.ncoeff, 1))
        _chk(lib().orc_assemble_vss(self.h, dptr(kd), dptr(kc)), "assemble")
        return kd, kc[:self.ncoeff]

    asym = False

    def set_asymmetric(self, on):
        _chk(lib().orc_set_asymmetric(self.h, int(bool(on))), "asymmetric")
        self.asym = bool(on)

    def csr_pattern(self):
        """The full CSR pattern of the equations (rows: the diagonal and
        every column coupled to it, ascending, 0-based) from the upper
        triangle of the sparsity pattern."""
        _, _, kp, ki = self.get_pattern()
        n = self.neqns
        rows = np.repeat(np.arange(n), kp)
        cols = ki.astype(np.int64) - 1
        r = np.concatenate([np.arange(n), rows, cols])
        c = np.concatenate([np.arange(n), cols, rows])
        order = np.lexsort((c, r))
        r, c = r[order], c[order]
        rp = np.zeros(n + 1, np.int64)
        np.cumsum(np.bincount(r, minlength=n), out=rp[1:])
        return rp, c.astype(np.int32)

    def assemble_csr(self, rp=None, col=None):
        """(rp, col, values) of the full CSR pattern (this model's own when
        rp is None) from the full element matrices of the last tangent
        (asymmetric assembly)."""
        if rp is None:
            rp, col = self.csr_pattern()
        rp = np.ascontiguousarray(rp, np.int64)
        col = np.ascontiguousarray(col, np.int32)
        val = np.zeros(max(int(rp[-1]), 1))
        _chk(lib().orc_assemble_csr(
            self.h, rp.ctypes.data_as(C.POINTER(C.c_int64)),
            col.ctypes.data_as(C.POINTER(C.c_int32)), dptr(val)),
            "assemble_csr")
        return rp, col, val[:int(rp[-1])]

    def get_ek_full(self, blk):
        """(span, T, T): element e's word [e, c, r] = B(r)^T D B(c)."""
        info = self.block_info(blk)
        span, T = info["span"], info["totdof"]
        out = np.zeros(span * T * T)
        _chk(lib().orc_get_ek_full(self.h, blk, dptr(out), out.size),
             "get_ek_full")
        return out.reshape(span, T, T)

    def commit_step(self):
        _chk(lib().orc_commit_step(self.h), "commit")

    def rollback(self):
        _chk(lib().orc_rollback(self.h), "rollback")

    def state_size(self, blk, which):
        i = self.block_info(blk)
        nq = i["span"] * i["ngp"]
        return {0: 9 * nq, 1: 9 * nq, 2: 6 * nq, 3: 6 * nq, 4: i["hist"] * nq,
                5: i["hist"] * nq, 6: 21 * nq, 7: i["utsz"] * i["span"],
                8: i["totdof"] * i["span"], 9: i["span"], 10: 9 * nq}[which]

    def get_block_state(self, blk, which):
        n = self.state_size(blk, which)
        out = np.zeros(n)
        _chk(lib().orc_get_block_state(self.h, blk, which, dptr(out), n),
             "get_block_state")
        return out

    def set_block_state(self, blk, which, data):
        data = np.ascontiguousarray(data, float)
        _chk(lib().orc_set_block_state(self.h, blk, which, dptr(data),
                                       data.size), "set_block_state")


REF_LEAF_PATH = os.path.join(_HERE, "_ref", "libw3d_refleaf.so")


def ref_leaf_built():
    """Whether `make -C oracle ref` has built the reference's leaf
    routines (without loading them)."""
    return os.path.exists(REF_LEAF_PATH)


def load_ref_leaf():
    """The reference's own leaf routines built by `make -C oracle ref`, or
    None when they are not built (no reference tree on this machine)."""
    if not ref_leaf_built():
        return None
    return C.CDLL(REF_LEAF_PATH, mode=os.RTLD_LAZY)
