"""MI355X-native FVCOM mode-split time step.

Host-side mirror of the reference's call surface: ``Model`` holds the grid
and state arrays under their ALL_VARS names (FVCOM_source/mod_main.F), and
``DeviceModel`` is the device-resident state whose ``internal_step`` and
``external_step`` replace INTERNAL_STEP (FVCOM_source/internal_step.F:29)
and EXTERNAL_STEP (FVCOM_source/external_step.F:30).  Both talk to the
C-ABI declared in ``include/fvgpu.h``; PyTorch is not needed here.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "lib")

# FVX extents in include/fvgpu.h order
_EXT = ["M1", "M2", "N1", "N1x3", "N1x4", "NE1", "NE1x2", "CV1", "CV1x2", "M1K", "N1K",
        "RSN", "RVE", "OBC", "OBCx2", "OBCxT", "TIDE", "RIV", "RIVxK", "RIVx2"]


class FvParams(C.Structure):
    _fields_ = [("dte", C.c_double), ("isplit", C.c_int32), ("iramp", C.c_int32),
                ("umol", C.c_double), ("vprnu", C.c_double), ("hprnu", C.c_double),
                ("cbcmin", C.c_double), ("min_depth", C.c_double),
                ("horizontal_closure", C.c_int32), ("vertical_closure", C.c_int32),
                ("scalar_positivity", C.c_int32), ("noflux_bot", C.c_int32),
                ("barotropic", C.c_int32), ("temperature_active", C.c_int32),
                ("salinity_active", C.c_int32), ("wind_on", C.c_int32),
                ("river_ts_specified", C.c_int32), ("type_tsobc", C.c_int32),
                ("obc_on", C.c_int32), ("obc_elevation_on", C.c_int32),
                ("iint0", C.c_int64), ("time0", C.c_double)]


class FvCase(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("dx", C.c_double), ("jitter", C.c_double),
                ("kb", C.c_int32), ("depth_kind", C.c_int32), ("n_seamounts", C.c_int32),
                ("ntide", C.c_int32), ("tide_amp", C.c_double * 3), ("river", C.c_int32),
                ("river_q", C.c_double), ("river_t", C.c_double), ("river_s", C.c_double),
                ("wind", C.c_int32), ("wind_mean", C.c_double), ("wind_std", C.c_double),
                ("wind_corr", C.c_double), ("ts_kind", C.c_int32), ("el_pert_std", C.c_double),
                ("el_pert_corr", C.c_double), ("coriolis", C.c_double), ("sponge_radius", C.c_double),
                ("sponge_coef", C.c_double), ("seed", C.c_uint64),
                ("p", FvParams)]


class FvModelHeader(C.Structure):
    """Leading scalar part of ``fv_model`` (arrays are reached by name)."""
    _fields_ = [(n, C.c_int32) for n in ("m", "n", "ne", "ncv", "ncv_i", "kb", "kbm1", "kbm2",
                                         "iobcn", "ntide", "numqbc", "nring_sn", "nring_ve")] + \
               [("p", FvParams)]


class FvDiag(C.Structure):
    _fields_ = [("nonfinite", C.c_int32), ("depth_fail", C.c_int32), ("min_depth", C.c_double),
                ("volume", C.c_double), ("salt", C.c_double), ("heat", C.c_double),
                ("iint", C.c_int64)]


_grid = None
_gpu = None


def lib_path(name):
    return os.path.join(_LIB, name)


def grid_lib():
    """Load libfvgrid.so (host metrics); raises if it has not been built."""
    global _grid
    if _grid is None:
        path = lib_path("libfvgrid.so")
        if not os.path.exists(path):
            raise RuntimeError("libfvgrid.so is not built; run myfvcomcode_amd.build.build_all()")
        L = C.CDLL(path)
        L.fv_last_error.restype = C.c_char_p
        L.fv_case_defaults.argtypes = [C.c_int32, C.POINTER(FvCase)]
        L.fv_model_build.argtypes = [C.POINTER(FvCase), C.POINTER(C.c_void_p)]
        L.fv_model_from_mesh.argtypes = [C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_int32, C.c_int32, C.c_void_p,
                                         C.POINTER(FvParams), C.POINTER(C.c_void_p)]
        L.fv_model_free.argtypes = [C.c_void_p]
        L.fv_model_field.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_void_p),
                                     C.POINTER(C.c_int64), C.POINTER(C.c_int32)]
        L.fv_model_nfields.restype = C.c_int32
        L.fv_model_field_name.restype = C.c_char_p
        L.fv_model_field_name.argtypes = [C.c_int32]
        L.fv_model_cold_start.argtypes = [C.c_void_p]
        L.fv_model_set_tide.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.fv_model_set_river.argtypes = [C.c_void_p, C.c_int32, C.c_double, C.c_double, C.c_double]
        L.fv_model_set_sponge.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.fv_model_set_obc_type.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
        L.fv_model_set_river_edge.argtypes = [C.c_void_p, C.c_int32, C.c_double, C.c_double, C.c_double]
        L.fv_rcm_order.argtypes = [C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.fv_model_union.argtypes = [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(C.c_void_p)]
        L.fv_extent_len.restype = C.c_int64
        L.fv_extent_len.argtypes = [C.c_void_p, C.c_int32]
        _grid = L
    return _grid


def gpu_lib():
    """Load libfvgpu.so; fails loudly when the HIP library is missing."""
    global _gpu
    if _gpu is None:
        path = lib_path("libfvgpu.so")
        if not os.path.exists(path):
            raise RuntimeError("libfvgpu.so (HIP) is not built; run myfvcomcode_amd.build.build_all()")
        L = C.CDLL(path)
        vp = C.c_void_p
        L.fvgpu_last_error.restype = C.c_char_p
        L.fv_last_error = L.fvgpu_last_error
        L.fvgpu_create.argtypes = [vp, C.c_int32, C.POINTER(vp)]
        L.fvgpu_internal_step.argtypes = [vp]
        L.fvgpu_run.argtypes = [vp, C.c_int64, C.c_int32]
        L.fvgpu_external_step.argtypes = [vp, C.c_int32]
        L.fvgpu_get_state.argtypes = [vp, C.c_char_p, vp, C.c_int64]
        L.fvgpu_put_state.argtypes = [vp, C.c_char_p, vp, C.c_int64]
        L.fvgpu_diagnostics.argtypes = [vp, C.POINTER(FvDiag)]
        L.fvgpu_kernel_time.argtypes = [vp, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        L.fvgpu_kernel_stat.argtypes = [vp, C.c_int32, C.POINTER(C.c_char_p), C.POINTER(C.c_double),
                                        C.POINTER(C.c_int64)]
        L.fvgpu_set_profiling.argtypes = [vp, C.c_int32]
        L.fvgpu_set_block.argtypes = [vp, C.c_char_p, C.c_int32]
        L.fvgpu_set_forcing.argtypes = [vp, C.c_char_p, C.c_void_p, C.c_int64]
        L.fvgpu_ensemble_update.argtypes = [vp, C.c_char_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_int64,
                                            C.c_int64, C.c_void_p]
        L.fvgpu_ensemble_update_noalloc.argtypes = [vp, C.c_char_p, C.c_int32, C.c_void_p, C.c_void_p,
                                                    C.c_int64, C.c_int64, C.c_void_p]
        L.fvgpu_set_concurrency.argtypes = [vp, C.c_int32]
        L.fvgpu_autotune.argtypes = [vp, C.c_int32, C.c_char_p, C.c_int64]
        fp = C.POINTER(C.c_float)
        L.fvgpu_set_closure.argtypes = [vp, C.c_int32, fp]
        L.fvgpu_closure_fallbacks.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.fvgpu_closure_samples.argtypes = [vp, fp, fp, C.c_int64, C.POINTER(C.c_int64)]
        L.fvmlp_nparams.restype = C.c_int32
        L.fvmlp_create.argtypes = [C.c_int64, fp, fp, C.c_int32, C.POINTER(vp)]
        L.fvmlp_set_params.argtypes = [vp, fp]
        L.fvmlp_get_params.argtypes = [vp, fp]
        L.fvmlp_grad.argtypes = [vp, C.POINTER(C.c_int32), C.c_int32, fp, fp]
        L.fvmlp_train.argtypes = [vp, C.c_int32, C.c_int32, C.c_int64, C.c_float, C.c_float, C.c_uint64, fp]
        L.fvmlp_infer.argtypes = [vp, C.c_int64, C.c_int64, C.c_int32, fp]
        L.fvmlp_destroy.argtypes = [vp]
        L.fvgpu_profile_only.argtypes = [vp, C.c_char_p]
        L.fvgpu_profile_every.argtypes = [vp, C.c_int32]
        L.fvgpu_synchronize.argtypes = [vp]
        L.fvgpu_stream_copy.argtypes = [vp, C.c_int64, C.c_int32, C.POINTER(C.c_double)]
        L.fvgpu_destroy.argtypes = [vp]
        L.fvgpu_device_count.argtypes = [C.POINTER(C.c_int32)]
        _gpu = L
    return _gpu


def _check(lib, rc, what):
    if rc != 0:
        msg = lib.fv_last_error()
        raise RuntimeError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else "?"))


def rcm_order(m, nv_rows):
    """Reverse Cuthill-McKee numbering (fv_rcm_order) of a mesh of m nodes
    and 1-based triangles nv_rows (N,3): (node_new, elem_new), the 1-based
    new index of each old node and element."""
    L = grid_lib()
    nv = np.ascontiguousarray(nv_rows, np.int32)
    node_new = np.zeros(m, np.int32)
    elem_new = np.zeros(nv.shape[0], np.int32)
    _check(L, L.fv_rcm_order(int(m), nv.shape[0], nv.ctypes.data, node_new.ctypes.data, elem_new.ctypes.data),
           "fv_rcm_order")
    return node_new, elem_new


def renumber_mesh(x, y, nv_rows, h, obc_nodes, node_new, elem_new):
    """The mesh in the numbering node_new / elem_new (1-based new index of
    each old one): (x, y, nv_rows, h, obc_nodes) ready for
    Model.from_mesh.  The open-boundary list keeps its order."""
    m = len(x)
    old_of_new = np.empty(m, np.int64)
    old_of_new[node_new - 1] = np.arange(m)
    nv = np.asarray(nv_rows)
    nv_new = np.empty_like(nv)
    nv_new[elem_new - 1] = node_new[nv - 1]
    return (np.asarray(x)[old_of_new], np.asarray(y)[old_of_new], nv_new, np.asarray(h)[old_of_new],
            node_new[np.asarray(obc_nodes, np.int64) - 1] if len(obc_nodes) else np.asarray(obc_nodes, np.int32))


class Model:
    """Host grid + metrics + forcing + state of one instance (libfvgrid)."""

    def __init__(self, ptr):
        self._ptr = C.c_void_p(ptr)
        self.hdr = FvModelHeader.from_address(ptr)

    @staticmethod
    def case_defaults(config):
        c = FvCase()
        L = grid_lib()
        _check(L, L.fv_case_defaults(int(config), C.byref(c)), "fv_case_defaults")
        return c

    @classmethod
    def build(cls, config=0, **overrides):
        """Synthetic configuration ``config`` (0, 1, 3) with field overrides
        (``nx=``, ``ny=``, ``kb=``, ``seed=`` ... or ``p_<param>=``)."""
        c = cls.case_defaults(config)
        for k, v in overrides.items():
            if k.startswith("p_"):
                setattr(c.p, k[2:], v)
            else:
                setattr(c, k, v)
        L = grid_lib()
        out = C.c_void_p()
        _check(L, L.fv_model_build(C.byref(c), C.byref(out)), "fv_model_build")
        return cls(out.value)

    @classmethod
    def from_mesh(cls, x, y, nv_rows, h, kb, obc_nodes=(), params=None):
        """Explicit mesh: clockwise 1-based triangles ``nv_rows`` (N,3)."""
        L = grid_lib()
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        h = np.ascontiguousarray(h, np.float64)
        nv = np.ascontiguousarray(nv_rows, np.int32)
        obc = np.ascontiguousarray(obc_nodes, np.int32)
        if params is None:
            params = cls.case_defaults(0).p
        out = C.c_void_p()
        rc = L.fv_model_from_mesh(len(x), nv.shape[0], x.ctypes.data, y.ctypes.data, nv.ctypes.data,
                                  h.ctypes.data, int(kb), len(obc), obc.ctypes.data if len(obc) else None,
                                  C.byref(params), C.byref(out))
        _check(L, rc, "fv_model_from_mesh")
        return cls(out.value)

    @classmethod
    def union(cls, parts):
        """Batched instances: the disjoint union of ``parts`` as one model.
        Returns (model, offsets) with offsets[p] = (node offset, element
        offset, node count, element count) of instance p."""
        L = grid_lib()
        arr = (C.c_void_p * len(parts))(*[p._ptr.value for p in parts])
        out = C.c_void_p()
        _check(L, L.fv_model_union(arr, len(parts), C.byref(out)), "fv_model_union")
        offs, om, on = [], 0, 0
        for p in parts:
            offs.append((om, on, p.m, p.n))
            om += p.m
            on += p.n
        return cls(out.value), offs

    @property
    def ptr(self):
        return self._ptr

    @property
    def m(self):
        return self.hdr.m

    @property
    def n(self):
        return self.hdr.n

    @property
    def kb(self):
        return self.hdr.kb

    def field(self, name):
        """Live numpy view of an ALL_VARS array (lower-case name)."""
        L = grid_lib()
        p = C.c_void_p()
        n = C.c_int64()
        dt = C.c_int32()
        _check(L, L.fv_model_field(self._ptr, name.encode(), C.byref(p), C.byref(n), C.byref(dt)),
               "field " + name)
        if n.value == 0 or not p.value:
            return np.zeros(0, np.float64 if dt.value == 0 else np.int32)
        ct = C.c_double if dt.value == 0 else C.c_int32
        arr = (ct * n.value).from_address(p.value)
        return np.ctypeslib.as_array(arr)

    def field3(self, name):
        """3-D array viewed as [KB, entity+1] (layer-major, like A(0:MT,KB))."""
        a = self.field(name)
        return a.reshape(self.kb, -1)

    def set_tide(self, periods, amps, phases_deg):
        """Spectral tide, uniform along the open boundary."""
        L = grid_lib()
        p = np.ascontiguousarray(periods, np.float64)
        a = np.ascontiguousarray(amps, np.float64)
        f = np.ascontiguousarray(phases_deg, np.float64)
        _check(L, L.fv_model_set_tide(self._ptr, len(p), p.ctypes.data, a.ctypes.data, f.ctypes.data),
               "fv_model_set_tide")

    def set_river(self, node, q, t=15.0, s=0.0):
        """Node river at a solid-boundary node (1-based index)."""
        L = grid_lib()
        _check(L, L.fv_model_set_river(self._ptr, int(node), float(q), float(t), float(s)),
               "fv_model_set_river")

    def set_river_edge(self, cell, q, t=15.0, s=0.0):
        """Edge river on a solid-boundary cell (1-based element index with
        exactly one solid edge)."""
        L = grid_lib()
        _check(L, L.fv_model_set_river_edge(self._ptr, int(cell), float(q), float(t), float(s)),
               "fv_model_set_river_edge")

    def set_sponge(self, nodes, radius, coef):
        """Sponge layer around nodes (1-based): CC_SPONGE = max(c*(1-dist/r))."""
        L = grid_lib()
        nd = np.ascontiguousarray(nodes, np.int32)
        r = np.ascontiguousarray(np.broadcast_to(radius, nd.shape), np.float64)
        c = np.ascontiguousarray(np.broadcast_to(coef, nd.shape), np.float64)
        _check(L, L.fv_model_set_sponge(self._ptr, len(nd), nd.ctypes.data, r.ctypes.data, c.ctypes.data),
               "fv_model_set_sponge")

    def set_obc_type(self, types):
        """TYPE_OBC of every open-boundary node (IOBCN values in 1..8): 1/2
        elevation specified, 3/4 zero elevation, 5/6 gravity-wave
        radiation, 7/8 Blumberg-Kantha; odd types linearise the node's
        adjacent cells (IUCP = 0)."""
        L = grid_lib()
        t = np.ascontiguousarray(types, np.int32)
        _check(L, L.fv_model_set_obc_type(self._ptr, len(t), t.ctypes.data), "fv_model_set_obc_type")

    def cold_start(self):
        """Recompute the cold-start state from the current T1/S1/EL fields."""
        L = grid_lib()
        _check(L, L.fv_model_cold_start(self._ptr), "fv_model_cold_start")

    def field_names(self):
        L = grid_lib()
        return [L.fv_model_field_name(i).decode() for i in range(L.fv_model_nfields())]

    def close(self):
        if self._ptr is not None and self._ptr.value:
            grid_lib().fv_model_free(self._ptr)
            self._ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


PRECISION = {"fp64": 0, "mixed": 1}


class DeviceModel:
    """Device-resident state of one model on the current GPU (libfvgpu)."""

    def __init__(self, model, precision="fp64"):
        L = gpu_lib()
        self.model = model
        h = C.c_void_p()
        _check(L, L.fvgpu_create(model.ptr, PRECISION[precision], C.byref(h)), "fvgpu_create")
        self._h = h

    def internal_step(self):
        L = gpu_lib()
        _check(L, L.fvgpu_internal_step(self._h), "fvgpu_internal_step")

    def run(self, nsteps, graph=True):
        L = gpu_lib()
        _check(L, L.fvgpu_run(self._h, int(nsteps), 1 if graph else 0), "fvgpu_run")

    def external_step(self, iext=1):
        L = gpu_lib()
        _check(L, L.fvgpu_external_step(self._h, int(iext)), "fvgpu_external_step")

    def get(self, name):
        L = gpu_lib()
        ref = self.model.field(name) if name in self.model.field_names() else None
        n = len(ref) if ref is not None else self._len_of(name)
        out = np.empty(n, np.float64)
        _check(L, L.fvgpu_get_state(self._h, name.encode(), out.ctypes.data, n), "get " + name)
        return out

    def _len_of(self, name):
        m1, n1, kb = self.model.m + 1, self.model.n + 1, self.model.kb
        node2 = {"elf", "dtfa", "egf"}
        elem2 = {"uard", "vard", "adx2d", "ady2d", "drx2d", "dry2d", "advua", "advva", "pstx", "psty",
                 "cbc", "wubot", "wvbot", "wusurf", "wvsurf", "wusurf2", "wvsurf2", "uaf", "vaf"}
        elem3 = {"uf", "vf", "advx", "advy", "drhox", "drhoy", "viscofm"}
        if name in node2:
            return m1
        if name in elem2:
            return n1
        if name in elem3:
            return n1 * kb
        return m1 * kb

    def put(self, name, values):
        L = gpu_lib()
        v = np.ascontiguousarray(values, np.float64)
        _check(L, L.fvgpu_put_state(self._h, name.encode(), v.ctypes.data, len(v)), "put " + name)

    def diagnostics(self):
        L = gpu_lib()
        d = FvDiag()
        _check(L, L.fvgpu_diagnostics(self._h, C.byref(d)), "fvgpu_diagnostics")
        return d

    def kernel_time(self, name):
        L = gpu_lib()
        us = C.c_double()
        cnt = C.c_int64()
        _check(L, L.fvgpu_kernel_time(self._h, name.encode(), C.byref(us), C.byref(cnt)),
               "fvgpu_kernel_time")
        return us.value, cnt.value

    def set_closure(self, mode, params=None):
        """0: MY2.5; 1: the learned closure with per-column MY2.5 fallback."""
        L = gpu_lib()
        p = None
        if params is not None:
            params = np.ascontiguousarray(params, np.float32)
            p = params.ctypes.data_as(C.POINTER(C.c_float))
        _check(L, L.fvgpu_set_closure(self._h, int(mode), p), "fvgpu_set_closure")

    def autotune(self, steps=1):
        """Pick the workgroup size of each kernel family by timing; the
        state is restored.  Returns {family: (size, microseconds)}."""
        L = gpu_lib()
        buf = C.create_string_buffer(8192)
        _check(L, L.fvgpu_autotune(self._h, int(steps), buf, len(buf)), "fvgpu_autotune")
        out = {}
        for item in buf.value.decode().split(";"):
            if item:
                f, rest = item.split("=")
                size, us = rest.split(":")
                out[f] = (int(size), float(us[:-2]))
        return out

    def set_block(self, family, size):
        _check(gpu_lib(), gpu_lib().fvgpu_set_block(self._h, family.encode(), int(size)), "fvgpu_set_block")

    def set_forcing(self, name, values):
        """Replace a forcing from the next step on: "wind" (UUWIND then
        VVWIND over elements 0..N, the wind at the end of the coming step;
        the previous one becomes its start and each external step takes
        the wind interpolated to its ExtTime), "qdis", "tdis" or "sdis"
        (per river)."""
        L = gpu_lib()
        v = np.ascontiguousarray(values, np.float64)
        _check(L, L.fvgpu_set_forcing(self._h, name.encode(), v.ctypes.data, len(v)), "set_forcing " + name)

    def set_concurrency(self, mode):
        """Independent chains of the step on a second stream, a bit mask:
        1 the T/S and Q2/Q2L pairs, 2 the 2-D stage pair (default 3, both),
        0 every chain on the main stream (per-kernel timing)."""
        _check(gpu_lib(), gpu_lib().fvgpu_set_concurrency(self._h, int(mode)), "fvgpu_set_concurrency")

    def closure_fallbacks(self):
        L = gpu_lib()
        nf, nc = C.c_int64(), C.c_int64()
        _check(L, L.fvgpu_closure_fallbacks(self._h, C.byref(nf), C.byref(nc)), "fvgpu_closure_fallbacks")
        return nf.value, nc.value

    def closure_samples(self):
        """(x[n, 8], y[n, 2]) features and MY2.5 KM/KH of the current state."""
        L = gpu_lib()
        n = C.c_int64()
        _check(L, L.fvgpu_closure_samples(self._h, None, None, 0, C.byref(n)), "fvgpu_closure_samples")
        x = np.zeros((n.value, 8), np.float32)
        y = np.zeros((n.value, 2), np.float32)
        fp = C.POINTER(C.c_float)
        _check(L, L.fvgpu_closure_samples(self._h, x.ctypes.data_as(fp), y.ctypes.data_as(fp), n.value,
                                          C.byref(n)), "fvgpu_closure_samples")
        return x, y

    def kernel_stats(self):
        """{family: (average us, launches)} of the events recorded so far."""
        L = gpu_lib()
        out = {}
        i = 0
        while True:
            nm = C.c_char_p()
            us = C.c_double()
            cnt = C.c_int64()
            if L.fvgpu_kernel_stat(self._h, i, C.byref(nm), C.byref(us), C.byref(cnt)) != 0:
                break
            out[nm.value.decode()] = (us.value, cnt.value)
            i += 1
        return out

    def set_profiling(self, on, kernel=None, every=1):
        """HIP events around launches (all families, or only ``kernel``),
        around one in ``every`` of them."""
        L = gpu_lib()
        _check(L, L.fvgpu_profile_only(self._h, (kernel or "").encode()), "fvgpu_profile_only")
        _check(L, L.fvgpu_profile_every(self._h, int(every)), "fvgpu_profile_every")
        _check(L, L.fvgpu_set_profiling(self._h, 1 if on else 0), "fvgpu_set_profiling")

    def stream_copy(self, n_doubles, reps=20):
        """Average microseconds of a plain copy of n doubles (k_copy)."""
        L = gpu_lib()
        us = C.c_double()
        _check(L, L.fvgpu_stream_copy(self._h, int(n_doubles), int(reps), C.byref(us)), "fvgpu_stream_copy")
        return us.value

    def synchronize(self):
        L = gpu_lib()
        _check(L, L.fvgpu_synchronize(self._h), "fvgpu_synchronize")

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            gpu_lib().fvgpu_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Trainer:
    """The closure MLP's device trainer (fvmlp): samples on the device,
    FP32 master weights, hand-written forward/backward/AdamW kernels."""

    def __init__(self, x, y, bmax=8192):
        L = gpu_lib()
        self.x = np.ascontiguousarray(x, np.float32)
        self.y = np.ascontiguousarray(y, np.float32)
        assert self.x.shape[1] == 8 and self.y.shape[1] == 2 and len(self.x) == len(self.y)
        fp = C.POINTER(C.c_float)
        h = C.c_void_p()
        _check(L, L.fvmlp_create(len(self.x), self.x.ctypes.data_as(fp), self.y.ctypes.data_as(fp), int(bmax),
                                 C.byref(h)), "fvmlp_create")
        self._h = h
        self.n = len(self.x)
        self.nparams = L.fvmlp_nparams()

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                gpu_lib().fvmlp_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparams
        _check(gpu_lib(), gpu_lib().fvmlp_set_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float))),
               "fvmlp_set_params")

    def params(self):
        p = np.zeros(self.nparams, np.float32)
        _check(gpu_lib(), gpu_lib().fvmlp_get_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float))),
               "fvmlp_get_params")
        return p

    def grad(self, idx=None, batch=None):
        L = gpu_lib()
        g = np.zeros(self.nparams, np.float32)
        loss = C.c_float()
        fp = C.POINTER(C.c_float)
        if idx is not None:
            idx = np.ascontiguousarray(idx, np.int32)
            rc = L.fvmlp_grad(self._h, idx.ctypes.data_as(C.POINTER(C.c_int32)), len(idx), g.ctypes.data_as(fp),
                              C.byref(loss))
        else:
            rc = L.fvmlp_grad(self._h, None, int(batch), g.ctypes.data_as(fp), C.byref(loss))
        _check(L, rc, "fvmlp_grad")
        return loss.value, g

    def train(self, steps, batch=8192, ntrain=None, lr=2e-3, wd=1e-4, seed=0):
        L = gpu_lib()
        loss = C.c_float()
        _check(L, L.fvmlp_train(self._h, int(steps), int(batch), int(ntrain or self.n), float(lr), float(wd),
                                int(seed), C.byref(loss)), "fvmlp_train")
        return loss.value

    def infer(self, lo=0, n=None, mfma=True):
        L = gpu_lib()
        n = int(n if n is not None else self.n - lo)
        y = np.zeros((n, 2), np.float32)
        _check(L, L.fvmlp_infer(self._h, int(lo), n, 1 if mfma else 0, y.ctypes.data_as(C.POINTER(C.c_float))),
               "fvmlp_infer")
        return y
