"""Host-side mirror of the reference's block interface over libw3dblk.

``BlockModel`` wraps one w3d_ctx: the calls follow the reference's per-pass
entry points (drive_eps_sig_internal_forces, tanstf, drive_assemble_solve,
update) and return numpy arrays in the reference's layouts.  Errors raise
``W3DError`` carrying the library's message; there is no fallback path: if
libw3dblk.so is missing or no HIP device is present, construction fails.
"""
import ctypes as C
import os

import numpy as np

from ._abi import (ASYM_PROTOS, LIB_BLOCK_PROTOS, PROTOS, SERVE_PROTOS,
                   W3D_CUT_STEP, W3DModel, dptr, iptr)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libw3dblk.so")
_LIB = None


class W3DError(RuntimeError):
    pass


def lib():
    """Load libw3dblk.so (built by warp3d_amd.build); raises if absent."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise W3DError(f"{LIB_PATH} is not built: run "
                           "python -m warp3d_amd.build")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in {**PROTOS, **SERVE_PROTOS, **ASYM_PROTOS,
                                  **LIB_BLOCK_PROTOS}.items():
            f = getattr(L, "w3d_" + name)
            f.restype = res
            f.argtypes = args
        L.w3d_create.restype = C.c_int32
        L.w3d_create.argtypes = [C.POINTER(C.c_void_p), C.POINTER(W3DModel),
                                 C.c_int32]
        L.w3d_destroy.argtypes = [C.c_void_p]
        L.w3d_last_error.restype = C.c_char_p
        L.w3d_version.restype = C.c_char_p
        L.w3d_tangent.argtypes = [C.c_void_p, C.c_int32, C.c_int32,
                                  C.c_int32]
        L.w3d_tangent.restype = C.c_int32
        L.w3d_recover_dev.argtypes = [C.c_void_p, C.c_int32, C.c_int32,
                                      C.c_int32, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.POINTER(C.c_double),
                                      C.POINTER(C.c_double)]
        L.w3d_recover_dev.restype = C.c_int32
        L.w3d_assemble_vss_dev.argtypes = [C.c_void_p, C.c_void_p,
                                           C.c_void_p]
        L.w3d_assemble_vss_dev.restype = C.c_int32
        L.w3d_synchronize.argtypes = [C.c_void_p]
        L.w3d_synchronize.restype = C.c_int32
        L.w3d_set_profiling.argtypes = [C.c_void_p, C.c_int32]
        L.w3d_set_profiling.restype = C.c_int32
        L.w3d_get_phase_time.argtypes = [C.c_void_p, C.c_int32,
                                         C.POINTER(C.c_int64),
                                         C.POINTER(C.c_double)]
        L.w3d_get_phase_time.restype = C.c_int32
        L.w3d_reset_phase_times.argtypes = [C.c_void_p]
        L.w3d_reset_phase_times.restype = C.c_int32
        L.w3d_set_time_increment.argtypes = [C.c_void_p, C.c_double]
        L.w3d_set_time_increment.restype = C.c_int32
        L.w3d_iteration_dev.argtypes = [C.c_void_p, C.c_int32, C.c_int32,
                                        C.c_int32, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_int32, C.c_void_p,
                                        C.c_void_p, C.POINTER(C.c_double),
                                        C.POINTER(C.c_double)]
        L.w3d_iteration_dev.restype = C.c_int32
        L.w3d_autotune.argtypes = [C.c_void_p, C.c_int32,
                                   C.POINTER(C.c_int32), C.c_int32,
                                   C.POINTER(C.c_int32)]
        L.w3d_autotune.restype = C.c_int32
        L.w3d_solve_pcg.argtypes = [C.c_void_p, C.POINTER(C.c_double),
                                    C.POINTER(C.c_double), C.c_double,
                                    C.c_int32, C.POINTER(C.c_int32),
                                    C.POINTER(C.c_double)]
        L.w3d_solve_pcg.restype = C.c_int32
        L.w3d_solve_pcg_dev.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_double,
                                        C.c_int32, C.POINTER(C.c_int32),
                                        C.POINTER(C.c_double)]
        L.w3d_solve_pcg_dev.restype = C.c_int32
        L.w3d_solve_pcg_static.argtypes = L.w3d_solve_pcg_dev.argtypes
        L.w3d_solve_pcg_static.restype = C.c_int32
        _LIB = L
    return _LIB


def _chk(rc, what):
    if rc < 0:
        raise W3DError(f"{what}: {lib().w3d_last_error().decode()}")
    return rc


class BlockModel:
    """A device-resident element-block model (one HIP device)."""

    def __init__(self, arrays, device=0):
        self._arrays = arrays
        self._m = arrays.struct()
        h = C.c_void_p()
        _chk(lib().w3d_create(C.byref(h), C.byref(self._m), device),
             "w3d_create")
        self.h = h
        self.nonode = arrays.nonode
        nb = C.c_int32()
        lib().w3d_get_blocking(self.h, C.byref(nb), None)
        self.nelblk = nb.value
        eb = np.zeros(4 * self.nelblk, np.int32)
        lib().w3d_get_blocking(self.h, C.byref(nb), iptr(eb))
        self.elblks = eb.reshape(self.nelblk, 4)
        self.neqns = None
        self.ncoeff = None

    def close(self):
        if getattr(self, "h", None):
            lib().w3d_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_time_increment(self, dt):
        _chk(lib().w3d_set_time_increment(self.h, dt),
             "w3d_set_time_increment")

    def set_temperatures(self, temper_nodes=None, dtemp_nodes=None,
                         temper_elems=None, dtemp_elems=None,
                         incremental=True):
        """The load step's temperatures (totals at n and the step's
        changes, nodal and element parts; None = zero)."""
        a = [None if v is None else np.ascontiguousarray(v, float)
             for v in (temper_nodes, dtemp_nodes, temper_elems, dtemp_elems)]
        _chk(lib().w3d_set_temperatures(
            self.h, *[dptr(v) if v is not None else None for v in a],
            int(bool(incremental))), "w3d_set_temperatures")

    def block_info(self, blk):
        info = np.zeros(5, np.int32)
        _chk(lib().w3d_block_info(self.h, blk, iptr(info)), "block_info")
        return dict(zip(("span", "ngp", "hist", "totdof", "utsz"),
                        info.tolist()))

    # -- drive_eps_sig_internal_forces --------------------------------
    def recover(self, step, iter_, u, du, extrapolated=False):
        ifv = np.zeros(3 * self.nonode)
        en = C.c_double()
        pw = C.c_double()
        rc = _chk(lib().w3d_recover(
            self.h, step, iter_, int(extrapolated),
            dptr(np.ascontiguousarray(u, float)),
            dptr(np.ascontiguousarray(du, float)), dptr(ifv),
            C.cast(C.byref(en), C.POINTER(C.c_double)),
            C.cast(C.byref(pw), C.POINTER(C.c_double))), "w3d_recover")
        return ifv, en.value, pw.value, rc

    def recover_dev(self, step, iter_, d_u, d_du, d_ifv, extrapolated=False):
        """Device-pointer form (integers from torch.Tensor.data_ptr())."""
        en = C.c_double()
        pw = C.c_double()
        rc = _chk(lib().w3d_recover_dev(self.h, step, iter_, int(extrapolated),
                                        d_u, d_du, d_ifv, C.byref(en),
                                        C.byref(pw)), "w3d_recover_dev")
        return en.value, pw.value, rc

    def recover_blocks(self, first, n, step, iter_, u, du,
                       extrapolated=False):
        return _chk(lib().w3d_recover_blocks(
            self.h, first, n, step, iter_, int(extrapolated),
            dptr(np.ascontiguousarray(u, float)),
            dptr(np.ascontiguousarray(du, float))), "w3d_recover_blocks")

    # -- tanstf ---------------------------------------------------------
    def tangent(self, step, iter_, precision=0):
        return _chk(lib().w3d_tangent(self.h, step, iter_, precision),
                    "w3d_tangent")

    def tangent_blocks(self, first, n, step, iter_, precision=0):
        return _chk(lib().w3d_tangent_blocks(self.h, first, n, step, iter_,
                                             precision), "w3d_tangent_blocks")

    # the per-block interface (do_nleps_block / do_nlek_block / addifv);
    # recover_block and tangent_block may run in several threads at once
    # for distinct blocks (ctypes releases the GIL during the call)
    def set_displacements(self, u, du):
        _chk(lib().w3d_set_displacements(
            self.h, dptr(np.ascontiguousarray(u, float)),
            dptr(np.ascontiguousarray(du, float))), "w3d_set_displacements")

    def recover_block(self, blk, step, iter_, extrapolated=False):
        """Returns (status, block energy, block plastic work, cut)."""
        en, pw, cut = C.c_double(), C.c_double(), C.c_int32()
        rc = _chk(lib().w3d_recover_block(
            self.h, blk, step, iter_, int(extrapolated),
            C.cast(C.byref(en), C.POINTER(C.c_double)),
            C.cast(C.byref(pw), C.POINTER(C.c_double)), C.byref(cut)),
            "w3d_recover_block")
        return rc, en.value, pw.value, cut.value

    def assemble_ifv(self):
        """Returns (ifv, sum_ifv, num_term_ifv)."""
        ifv = np.zeros(3 * self.nonode)
        s, n = C.c_double(), C.c_int64()
        _chk(lib().w3d_assemble_ifv(
            self.h, dptr(ifv), C.cast(C.byref(s), C.POINTER(C.c_double)),
            C.byref(n)), "w3d_assemble_ifv")
        return ifv, s.value, n.value

    def ifv_sums(self):
        """addifv's (sum_ifv, num_term_ifv) of the last whole-pass
        recovery that returned its energy."""
        s, n = C.c_double(), C.c_int64()
        _chk(lib().w3d_get_ifv_sums(
            self.h, C.cast(C.byref(s), C.POINTER(C.c_double)), C.byref(n)),
            "w3d_get_ifv_sums")
        return s.value, n.value

    def tangent_block(self, blk, step, iter_, precision=0):
        return _chk(lib().w3d_tangent_block(self.h, blk, step, iter_,
                                            precision), "w3d_tangent_block")

    # -- assembly --------------------------------------------------------
    def pattern(self, cstmap):
        neq = C.c_int32()
        nco = C.c_int64()
        _chk(lib().w3d_pattern(self.h, iptr(np.ascontiguousarray(
            cstmap, np.int32)), C.byref(neq), C.byref(nco)), "w3d_pattern")
        self.neqns, self.ncoeff = neq.value, nco.value
        return self.neqns, self.ncoeff

    def get_pattern(self):
        d = np.zeros(3 * self.nonode, np.int32)
        e = np.zeros(self.neqns, np.int32)
        p = np.zeros(self.neqns, np.int32)
        k = np.zeros(max(self.ncoeff, 1), np.int32)
        _chk(lib().w3d_get_pattern(self.h, iptr(d), iptr(e), iptr(p),
                                   iptr(k)), "w3d_get_pattern")
        return d, e, p, k[:self.ncoeff]

    def assemble_vss(self):
        kd = np.zeros(self.neqns or 0)
        kc = np.zeros(max(self.ncoeff or 0, 1))
        _chk(lib().w3d_assemble_vss(self.h, dptr(kd), dptr(kc)),
             "w3d_assemble_vss")
        return kd, kc[:self.ncoeff]

    asym = False

    def set_asymmetric(self, on):
        """Asymmetric assembly (w3d_set_asymmetric)."""
        _chk(lib().w3d_set_asymmetric(self.h, int(bool(on))),
             "w3d_set_asymmetric")
        self.asym = bool(on)

    def csr_pattern(self):
        nnz = C.c_int64()
        _chk(lib().w3d_get_csr_pattern(self.h, C.byref(nnz), None, None),
             "w3d_get_csr_pattern")
        rp = np.zeros(self.neqns + 1, np.int64)
        col = np.zeros(max(nnz.value, 1), np.int32)
        _chk(lib().w3d_get_csr_pattern(
            self.h, None, rp.ctypes.data_as(C.POINTER(C.c_int64)),
            col.ctypes.data_as(C.POINTER(C.c_int32))), "w3d_get_csr_pattern")
        return rp, col[:nnz.value]

    def assemble_csr(self):
        """(rp, col, values) of the full CSR rows (asymmetric assembly)."""
        rp, col = self.csr_pattern()
        val = np.zeros(max(col.size, 1))
        _chk(lib().w3d_assemble_csr(self.h, dptr(val)), "w3d_assemble_csr")
        return rp, col, val[:col.size]

    def solve_bicgstab(self, rhs, rtol=1e-12, maxit=20000):
        """x of the asymmetric CSR of the last assemble_csr (BiCGStab);
        returns (x, iterations, relative residual)."""
        x = np.zeros(self.neqns)
        it = C.c_int32()
        rr = C.c_double()
        _chk(lib().w3d_solve_bicgstab(
            self.h, dptr(np.ascontiguousarray(rhs, float)), dptr(x), rtol,
            maxit, C.byref(it), C.byref(rr)), "w3d_solve_bicgstab")
        return x, it.value, rr.value

    def assemble_csr_dev(self, d_val):
        return _chk(lib().w3d_assemble_csr_dev(self.h, d_val),
                    "w3d_assemble_csr_dev")

    def get_ek_full(self, blk):
        """(span, T, T): element e's word [e, c, r] = B(r)^T D B(c)."""
        info = self.block_info(blk)
        span, T = info["span"], info["totdof"]
        out = np.zeros(span * T * T)
        _chk(lib().w3d_get_block_ek_full(self.h, blk, dptr(out), out.size),
             "w3d_get_block_ek_full")
        return out.reshape(span, T, T)

    def assemble_vss_dev(self, d_kdiag, d_kcoeffs):
        return _chk(lib().w3d_assemble_vss_dev(self.h, d_kdiag, d_kcoeffs),
                    "w3d_assemble_vss_dev")

    def iteration_dev(self, step, iter_, d_u, d_du, d_ifv, precision,
                      d_kdiag, d_kcoeffs, extrapolated=False):
        """One graph-launched Newton iteration on device buffers
        (recovery, tangent and assembly); returns (energy, plastic work,
        status)."""
        en = C.c_double()
        pw = C.c_double()
        rc = _chk(lib().w3d_iteration_dev(
            self.h, step, iter_, int(extrapolated), d_u, d_du, d_ifv,
            precision, d_kdiag, d_kcoeffs, C.byref(en), C.byref(pw)),
            "w3d_iteration_dev")
        return en.value, pw.value, rc

    def autotune(self, precision=0):
        """Runs the autotuner; returns {group key: elements per work item}
        and, with mm10 blocks, {-1: mm10 workgroups}."""
        buf = (C.c_int32 * 64)()
        n = C.c_int32()
        _chk(lib().w3d_autotune(self.h, precision, buf, 64, C.byref(n)),
             "w3d_autotune")
        v = list(buf)[:n.value]
        return dict(zip(v[0::2], v[1::2]))

    # -- linear solve (Jacobi-PCG on the last assembly) ------------------
    def solve_pcg(self, rhs, x0=None, rtol=1e-10, maxit=10000):
        """Host vectors; returns (x, iterations, relative residual)."""
        x = np.zeros(self.neqns) if x0 is None else np.array(x0, float)
        it = C.c_int32()
        rr = C.c_double()
        _chk(lib().w3d_solve_pcg(self.h, dptr(np.ascontiguousarray(rhs,
                                                                   float)),
                                 dptr(x), rtol, maxit, C.byref(it),
                                 C.byref(rr)), "w3d_solve_pcg")
        return x, it.value, rr.value

    def solve_pcg_dev(self, d_kdiag, d_kcoeffs, d_rhs, d_x, rtol=1e-10,
                      maxit=10000):
        """Device vectors (d_x: initial guess in, solution out); returns
        (iterations, relative residual)."""
        it = C.c_int32()
        rr = C.c_double()
        _chk(lib().w3d_solve_pcg_dev(self.h, d_kdiag, d_kcoeffs, d_rhs, d_x,
                                     rtol, maxit, C.byref(it), C.byref(rr)),
             "w3d_solve_pcg_dev")
        return it.value, rr.value

    def solve_pcg_static(self, d_kdiag, d_kcoeffs, d_rhs, d_x, rtol=1e-10,
                         maxit=10000):
        """solve_pcg_dev with the library's static work memory (no
        device allocation)."""
        it = C.c_int32()
        rr = C.c_double()
        _chk(lib().w3d_solve_pcg_static(self.h, d_kdiag, d_kcoeffs, d_rhs,
                                        d_x, rtol, maxit, C.byref(it),
                                        C.byref(rr)), "w3d_solve_pcg_static")
        return it.value, rr.value

    # -- state -----------------------------------------------------------
    def commit_step(self):
        _chk(lib().w3d_commit_step(self.h), "w3d_commit_step")

    def rollback(self):
        _chk(lib().w3d_rollback(self.h), "w3d_rollback")

    def synchronize(self):
        _chk(lib().w3d_synchronize(self.h), "w3d_synchronize")

    # -- phase timing (HIP events on the library's stream) ----------------
    PHASES = ("recover", "ifv_gather", "energy_reduce", "tangent",
              "assemble", "mm10")

    def set_profiling(self, on=True):
        _chk(lib().w3d_set_profiling(self.h, int(bool(on))),
             "w3d_set_profiling")

    def reset_phase_times(self):
        _chk(lib().w3d_reset_phase_times(self.h), "w3d_reset_phase_times")

    def phase_times(self):
        out = {}
        for k, name in enumerate(self.PHASES):
            n = C.c_int64()
            ms = C.c_double()
            _chk(lib().w3d_get_phase_time(self.h, k, C.byref(n),
                                          C.byref(ms)), "phase_time")
            out[name] = (n.value, ms.value)
        return out

    def state_size(self, blk, which):
        i = self.block_info(blk)
        nq = i["span"] * i["ngp"]
        return {0: 9 * nq, 1: 9 * nq, 2: 6 * nq, 3: 6 * nq, 4: i["hist"] * nq,
                5: i["hist"] * nq, 6: 21 * nq, 7: i["utsz"] * i["span"],
                8: i["totdof"] * i["span"], 9: i["span"], 10: 9 * nq}[which]

    def get_block_state(self, blk, which):
        n = self.state_size(blk, which)
        out = np.zeros(n)
        _chk(lib().w3d_get_block_state(self.h, blk, which, dptr(out), n),
             "w3d_get_block_state")
        return out

    def set_block_state(self, blk, which, data):
        data = np.ascontiguousarray(data, float)
        _chk(lib().w3d_set_block_state(self.h, blk, which, dptr(data),
                                       data.size), "w3d_set_block_state")


__all__ = ["BlockModel", "W3DError", "lib", "W3D_CUT_STEP"]
