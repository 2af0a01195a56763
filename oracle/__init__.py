"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of the serial FP64 oracle.

The oracle restates FVCOM's INTERNAL_STEP / EXTERNAL_STEP
(FVCOM_source/internal_step.F:29, external_step.F:30) in the reference's
scatter order (oracle/fvcom_oracle.c).  Only tests/, the smoke check and
the benchmark's CPU leg may load it; the product never does.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "libfvoracle.so")
_SO_NATIVE = os.path.join(_HERE, "_build", "libfvoracle_native.so")
_libs = {}


def build(verbose=False, native=False):
    """Build the oracle; native=True builds the copy tuned for this host
    (-march=native) that the CPU baseline times."""
    subprocess.run(["make", "-s", "-C", _HERE] + (["native"] if native else []), check=True,
                   stdout=None if verbose else subprocess.DEVNULL)
    return _SO_NATIVE if native else _SO


def lib(native=False):
    L = _libs.get(native)
    if L is None:
        so = _SO_NATIVE if native else _SO
        if not os.path.exists(so):
            build(native=native)
        L = C.CDLL(so)
        vp = C.c_void_p
        L.orc_create.restype = vp
        L.orc_create.argtypes = [vp]
        L.orc_destroy.argtypes = [vp]
        L.orc_internal_step.argtypes = [vp]
        L.orc_run.argtypes = [vp, C.c_longlong]
        L.orc_external_only.argtypes = [vp, C.c_int]
        L.orc_get.argtypes = [vp, C.c_char_p, vp, C.c_longlong]
        L.orc_put.argtypes = [vp, C.c_char_p, vp, C.c_longlong]
        L.orc_iint.restype = C.c_longlong
        L.orc_iint.argtypes = [vp]
        L.orc_depth_fail.argtypes = [vp]
        L.orc_routine.argtypes = [vp, C.c_char_p, C.c_int]
        L.orc_budgets.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        _libs[native] = L
    return L


class Oracle:
    """FP64 CPU state of one model, stepped by the oracle."""

    def __init__(self, model, native=False):
        self.model = model  # keep the grid alive
        self._L = lib(native)
        ptr = self._L.orc_create(model.ptr)
        if not ptr:
            raise ValueError("the oracle takes 2 <= KB <= 64 (KB = %d)" % model.kb)
        self._o = C.c_void_p(ptr)

    def internal_step(self):
        self._L.orc_internal_step(self._o)

    def run(self, nsteps):
        self._L.orc_run(self._o, int(nsteps))

    def external_step(self, iext=1):
        self._L.orc_external_only(self._o, int(iext))

    def _len(self, name):
        m1, n1, kb = self.model.m + 1, self.model.n + 1, self.model.kb
        node2 = {"el", "et", "d", "dt", "elf", "dtfa", "egf", "elrk", "elm1", "elm2"}
        elem2 = {"ua", "va", "el1", "et1", "d1", "dt1", "uard", "vard", "adx2d", "ady2d", "drx2d",
                 "dry2d", "advua", "advva", "pstx", "psty", "cbc", "wubot", "wvbot", "wusurf",
                 "wvsurf", "wusurf2", "wvsurf2", "uaf", "vaf", "uark", "vark", "elrk1", "elf1"}
        elem3 = {"u", "v", "w", "ww", "uf", "vf", "advx", "advy", "drhox", "drhoy", "rho", "rmean",
                 "t", "s", "km1", "viscofm"}
        if name == "ramp":
            return 1
        if name == "uard_obcn":
            return self.model.hdr.iobcn
        if name == "xflux_obc":
            return self.model.hdr.iobcn * kb
        if name in node2:
            return m1
        if name in elem2:
            return n1
        if name in elem3:
            return n1 * kb
        return m1 * kb

    def get(self, name):
        n = self._len(name)
        out = np.empty(n, np.float64)
        if self._L.orc_get(self._o, name.encode(), out.ctypes.data, n) != 0:
            raise KeyError(name)
        return out

    def put(self, name, values):
        v = np.ascontiguousarray(values, np.float64)
        if self._L.orc_put(self._o, name.encode(), v.ctypes.data, len(v)) != 0:
            raise KeyError(name)

    def routine(self, name, arg=0):
        """Run one routine of the step on the current state (viscof_h,
        wreal, vdif_q, bcond_ts with NCON = arg, adjust2d3d of type arg,
        bottom_roughness, baropg, vdif_uv, advave (ADVUA/ADVVA), advection
        (ADVX/ADVY), adv_uv (UF/VF), adv_q (Q2F from Q2), adv_ts (TF1 or SF1
        for NCON = arg), extuv (EXTUV_EDGE of RK stage arg), vertvl,
        fct_ts (FCT_T for arg 1, FCT_S for 2), fct_q (FCT_Q2 for arg 1,
        FCT_Q2L for 2), adjust_ts, bcond_gcn2 on UAF/VAF, bcond_gcn5 on
        U/V)."""
        if self._L.orc_routine(self._o, name.encode(), int(arg)) != 0:
            raise KeyError(name)

    @property
    def iint(self):
        return self._L.orc_iint(self._o)

    @property
    def depth_fail(self):
        return bool(self._L.orc_depth_fail(self._o))

    def budgets(self):
        v, s, h = C.c_double(), C.c_double(), C.c_double()
        self._L.orc_budgets(self._o, C.byref(v), C.byref(s), C.byref(h))
        return v.value, s.value, h.value

    def close(self):
        if self._o is not None and self._o.value:
            self._L.orc_destroy(self._o)
            self._o = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
