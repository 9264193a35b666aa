"""MI355X back-end of the Fletcher TTI propagation path.

Python mirror of the C ABI in include/fletcher/*.h (loaded with ctypes from the
in-tree libfletcher_amd.so; nothing here computes on the CPU except host model
setup, which is the C library's too).

    Driver      the reference plug-in API (original/driver.h:8-22): initialize,
                insert_source, propagate, update_pointers, finalize, plus the
                fl_hip.h extensions (variants, precision, tiles, traces, graph runs)
    Batch       many shots advanced by one launch per step over (shot, tile) items
    Ref64       the FP64 device reference (der1der1 in double, exact weights)
    model_*     host model setup (main.c, boundary.c, precomp.h restated)

Arrays are numpy float32 volumes shaped (sz, sy, sx), x fastest (map.h:14).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# FL_LIBRARY selects another build of the same C ABI (A/B comparisons of kernel builds)
LIB_PATH = os.environ.get("FL_LIBRARY") or os.path.join(_HERE, "libfletcher_amd.so")
CLI_PATH = os.path.join(_HERE, "bin", "fletcher")

BORD = 4  # stencil halo (main.c:21)
FORM = {"ISO": 0, "VTI": 1, "TTI": 2}
VARIANT = {"fused": 0, "direct": 1, "twopass": 2}
PRECISION = {"fp32": 0, "bf16": 1, "fp16": 2, "fp16r": 3, "fp32r": 4, "q16r": 5}
FIELD = {"pp": 0, "pc": 1, "qp": 2, "qc": 3}
COEF_NAMES = ("ch1dxx", "ch1dyy", "ch1dzz", "ch1dxy", "ch1dyz", "ch1dxz",
              "v2px", "v2pz", "v2sz", "v2pn")
MODEL_NAMES = ("vpz", "vsv", "epsilon", "delta", "phi", "theta")


class FletcherError(RuntimeError):
    pass


_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_lib = None


def lib():
    """The loaded C-ABI library; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FletcherError("%s is missing: build it with `python -m fletcher_petrobras_amd.build`"
                            % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    i, f, d, ll, vp = C.c_int, C.c_float, C.c_double, C.c_longlong, C.c_void_p
    sig = {
        "DRIVER_Initialize": (None, [i, i, i, i, f, f, f, f] + [_f32p] * 6 + [vp] * 4),
        "DRIVER_Finalize": (None, []),
        "DRIVER_Propagate": (None, [i, i, i, i, f, f, f, f, i, vp, vp, vp, vp]),
        "DRIVER_Update_pointers": (None, [i, i, i, vp]),
        "DRIVER_InsertSource": (None, [f, i, i, vp, vp, f]),
        "fl_last_error": (C.c_char_p, []),
        "fl_status": (i, []),
        "fl_clear_error": (None, []),
        "fl_set_variant": (i, [i]),
        "fl_get_variant": (i, []),
        "fl_set_precision": (i, [i]),
        "fl_get_precision": (i, []),
        "fl_set_sync": (i, [i]),
        "fl_set_tiles": (i, [i, i]),
        "fl_set_stream_hints": (i, [i]),
        "fl_set_tile_strip": (i, [i]),
        "fl_batch_set_persistent": (i, [i]),
        "fl_get_tile_strip": (i, []),
        "fl_set_lane_pairs": (i, [i]),
        "fl_get_lane_pairs": (i, []),
        "fl_get_stream_hints": (i, []),
        "fl_get_tiles": (i, [C.POINTER(i), C.POINTER(i)]),
        "fl_autotune": (i, [i, C.c_char_p]),
        "fl_get_field": (i, [i, _f32p]),
        "fl_set_field": (i, [i, _f32p]),
        "fl_record_traces": (i, [i, i, i]),
        "fl_get_traces": (i, [_f32p, i]),
        "fl_run_steps": (i, [i, i, i, _f32p, i]),
        "fl_reset_fields": (i, []),
        "fl_initialize_streamed": (i, [i, i, i, i, f, f, f, f, vp, vp]),
        "fl_set_device_precompute": (i, [i]),
        "fl_run_steps_static": (i, [i, i, i, _f32p]),
        "fl_synchronize": (i, []),
        "fl_set_timing": (i, [i]),
        "fl_kernel_times": (i, [C.POINTER(d), C.POINTER(i)]),
        "fl_region_begin": (i, []),
        "fl_region_end": (i, [C.POINTER(d)]),
        "fl_device_name": (i, [C.c_char_p, i]),
        "fl_get_derivative_volume": (i, [i, _f32p]),
        "fl_record_bytes": (ll, []),
        "fl_get_records": (i, [vp]),
        "fl_device_bytes": (ll, []),
        "fl_device_memory": (i, [C.POINTER(ll), C.POINTER(ll)]),
        "fl_bytes_per_cell": (i, []),
        "fl_batch_create": (i, [i, vp, i, f, f, f, f, i]),
        "fl_batch_set_tiles": (i, [i, i]),
        "fl_batch_get_tiles": (i, [C.c_void_p, C.c_void_p]),
        "fl_batch_run": (i, [i, i, _f32p, i, i, i]),
        "fl_batch_get_field": (i, [i, i, _f32p]),
        "fl_batch_get_traces": (i, [i, _f32p, i]),
        "fl_batch_autotune": (i, [i, C.c_char_p]),
        "fl_batch_kernel_ms": (i, [C.POINTER(d)]),
        "fl_batch_reset": (i, []),
        "fl_batch_destroy": (None, []),
        "fl_ref64_create": (i, [i, i, i, i, f, f, f, f] + [_f32p] * 6),
        "fl_ref64_run": (i, [i, i, i, _f32p, i, i, i]),
        "fl_ref64_get_field": (i, [i, _f64p]),
        "fl_ref64_get_traces": (i, [_f64p, i]),
        "fl_ref64_destroy": (None, []),
        # host model setup (fl_model.h)
        "fl_rand_seed": (None, [vp, C.c_uint32]),
        "fl_rand_next": (C.c_int32, [vp]),
        "fl_model_homogeneous": (i, [i, C.c_int64] + [_f32p] * 6),
        "fl_recommended_dt": (f, [C.c_int64, _f32p, _f32p, f, f, f, d]),
        "fl_random_boundary": (None, [C.c_uint32] + [i] * 8 + [_f32p, _f32p]),
        "fl_source": (f, [f, i]),
        "fl_precompute_f32": (None, [C.c_int64] + [_f32p] * 6 + [C.POINTER(vp)]),
        "fl_precompute_f64": (None, [C.c_int64] + [_f32p] * 6 + [C.POINTER(vp)]),
        "fl_num_steps": (i, [f, f]),
        "fl_model_gradient": (i, [vp] + [i] * 8 + [_f32p] * 6),
        "fl_gradient_model_defaults": (None, [vp, C.c_uint64]),
        "fl_model_layered": (i, [vp] + [i] * 8 + [_f32p] * 6),
        "fl_layered_model_defaults": (None, [vp, C.c_uint64]),
        "fl_cfl_dt": (f, [C.c_int64, _f32p, _f32p, f, f]),
        "fl_model_stream_gradient": (vp, [vp] + [i] * 8),
        "fl_model_stream_layered": (vp, [vp] + [i] * 8),
        "fl_model_stream_fill": (i, [vp, i, i] + [_f32p] * 6),
        "fl_model_stream_cfl_dt": (f, [vp, f, f]),
        "fl_model_stream_close": (None, [vp]),
        "fl_f64_to_f16": (C.c_uint16, [d]),
        "fl_f64_to_bf16": (C.c_uint16, [d]),
        "fl_angle_u16": (C.c_uint16, [d]),
        "fl_pack_coefs16": (None, [C.c_int64] + [_f32p] * 6 + [d, i, C.POINTER(vp)]),
        "fl_pack_recompute16": (None, [C.c_int64] + [_f32p] * 6 + [d, C.POINTER(vp)]),
        "fl_pack_recompute32": (None, [C.c_int64] + [_f32p] * 6 + [C.POINTER(vp)]),
        "fl_recompute_q16_base": (i, [C.c_int64] + [_f32p] * 4 + [d]),
        "fl_pack_recompute_q16": (None, [C.c_int64] + [_f32p] * 6 + [d, i, C.POINTER(vp)]),
        "fl_extract_window": (None, [_f32p] + [i] * 9 + [_f32p]),
        "fl_rsf_write_field": (i, [C.c_char_p] + [i] * 9 + [f, f, f, _f32p]),
        "fl_rsf_write_traces": (i, [C.c_char_p, i, i, i, f, f, f, _f32p]),
        "fl_rsf_open": (vp, [C.c_char_p] + [i] * 6 + [f] * 4),
        "fl_rsf_append": (i, [vp, i, i, i, _f32p]),
        "fl_rsf_close": (i, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _check(rc):
    if rc is not None and rc < 0:
        raise FletcherError(lib().fl_last_error().decode())
    return rc


def last_error():
    return lib().fl_last_error().decode()


def _c32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptrs(arrs):
    return (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


# ------------------------------------------------------------------ host model setup

def grid_shape(nx, ny, nz, absorb, bord=BORD):
    """Array extents (sz, sy, sx) = n + 2 bord + 2 absorb (main.c:95-97)."""
    return (nz + 2 * bord + 2 * absorb, ny + 2 * bord + 2 * absorb, nx + 2 * bord + 2 * absorb)


def index(shape, ix, iy, iz):
    sz, sy, sx = shape
    return (iz * sy + iy) * sx + ix


def center_index(shape):
    """iSource of main.c:105-108."""
    sz, sy, sx = shape
    return index(shape, sx // 2, sy // 2, sz // 2)


def model_homogeneous(form, shape):
    """ISO/VTI/TTI model of main.c:168-221 as (vpz, vsv, epsilon, delta, phi, theta)."""
    m = [np.empty(shape, np.float32) for _ in range(6)]
    rc = lib().fl_model_homogeneous(FORM[form] if isinstance(form, str) else form, m[0].size, *m)
    if rc != 0:
        raise FletcherError("unknown formulation %r" % (form,))
    return m


def random_boundary(model, n3, absorb, seed=1, bord=BORD):
    """Apply boundary.c:7-98 to vpz, vsv of `model` in place (rand() stream srand(seed))."""
    vpz, vsv = model[0], model[1]
    sz, sy, sx = vpz.shape
    lib().fl_random_boundary(seed, sx, sy, sz, n3[0], n3[1], n3[2], bord, absorb, vpz, vsv)
    return model


def recommended_dt(model, d, mi=0.2):
    return lib().fl_recommended_dt(model[0].size, model[0], model[2], d[0], d[1], d[2], mi)


def num_steps(tmax, dt):
    return lib().fl_num_steps(tmax, dt)


def source(dt, it):
    """Ricker sample at t = it*dt (source.c:7-15)."""
    return lib().fl_source(dt, it)


def sources(dt, n, it0=0):
    """src[k] = Source(dt, it0+k): the value inserted before step it0+1+k."""
    return np.array([lib().fl_source(dt, it0 + k) for k in range(n)], dtype=np.float32)


def precompute(model, fp="f32"):
    """The 10 coefficient volumes of precomp.h (order COEF_NAMES)."""
    dt = np.float32 if fp == "f32" else np.float64
    out = [np.empty(model[0].shape, dt) for _ in range(10)]
    fn = lib().fl_precompute_f32 if fp == "f32" else lib().fl_precompute_f64
    fn(model[0].size, *[_c32(m) for m in model], _ptrs(out))
    return out


class _Gradient(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("vmin", C.c_float), ("vmax", C.c_float), ("nblobs", C.c_int),
                ("blob_amp", C.c_float), ("blob_smin", C.c_float), ("blob_smax", C.c_float),
                ("eps_max", C.c_float), ("delta_max", C.c_float), ("theta_max", C.c_float),
                ("phi_max", C.c_float), ("nsmooth", C.c_int), ("smooth_smin", C.c_float),
                ("smooth_smax", C.c_float), ("boundary_seed", C.c_uint32)]


class _Layered(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("nlayers", C.c_int), ("vmin", C.c_float), ("vmax", C.c_float),
                ("max_dip_deg", C.c_float), ("eps_max", C.c_float), ("delta_max", C.c_float),
                ("theta_max", C.c_float), ("phi_max", C.c_float), ("nsmooth", C.c_int),
                ("smooth_smin", C.c_float), ("smooth_smax", C.c_float), ("boundary_seed", C.c_uint32)]


def model_gradient(n3, absorb, seed=344, bord=BORD, **overrides):
    """Seeded heterogeneous TTI model: vpz gradient + Gaussian blobs, smooth eps/delta/theta/phi,
    reference random boundary.  n3 = (nx, ny, nz) interior."""
    p = _Gradient()
    lib().fl_gradient_model_defaults(C.byref(p), seed)
    for k, v in overrides.items():
        setattr(p, k, v)
    shape = grid_shape(n3[0], n3[1], n3[2], absorb, bord)
    sz, sy, sx = shape
    m = [np.empty(shape, np.float32) for _ in range(6)]
    _check(lib().fl_model_gradient(C.byref(p), sx, sy, sz, n3[0], n3[1], n3[2], bord, absorb, *m))
    return m


def model_layered(n3, absorb, seed=1024, bord=BORD, **overrides):
    """Seeded dipping-layer TTI model with smooth theta/phi and the reference boundary."""
    p = _Layered()
    lib().fl_layered_model_defaults(C.byref(p), seed)
    for k, v in overrides.items():
        setattr(p, k, v)
    shape = grid_shape(n3[0], n3[1], n3[2], absorb, bord)
    sz, sy, sx = shape
    m = [np.empty(shape, np.float32) for _ in range(6)]
    _check(lib().fl_model_layered(C.byref(p), sx, sy, sz, n3[0], n3[1], n3[2], bord, absorb, *m))
    return m


def device_memory():
    """(free, total) bytes of the device."""
    f, t = C.c_longlong(), C.c_longlong()
    _check(lib().fl_device_memory(C.byref(f), C.byref(t)))
    return f.value, t.value


class ModelStream:
    """A seeded gradient or layered model produced in z slabs (fl_model_stream_*): the planes
    of model_gradient / model_layered without the full arrays on the host."""

    def __init__(self, kind, n3, absorb, seed, bord=BORD, **overrides):
        L = lib()
        if kind == "gradient":
            p = _Gradient()
            L.fl_gradient_model_defaults(C.byref(p), seed)
        else:
            p = _Layered()
            L.fl_layered_model_defaults(C.byref(p), seed)
        for k, v in overrides.items():
            setattr(p, k, v)
        self.shape = grid_shape(n3[0], n3[1], n3[2], absorb, bord)
        sz, sy, sx = self.shape
        opener = L.fl_model_stream_gradient if kind == "gradient" else L.fl_model_stream_layered
        self.handle = opener(C.byref(p), sx, sy, sz, n3[0], n3[1], n3[2], bord, absorb)
        if not self.handle:
            raise FletcherError("model stream: host allocation failed")

    def fill(self, z0, nz):
        """Planes z0 .. z0+nz-1 of the six arrays (slabs in increasing z from plane 0)."""
        sz, sy, sx = self.shape
        out = [np.empty((nz, sy, sx), np.float32) for _ in range(6)]
        _check(lib().fl_model_stream_fill(self.handle, z0, nz, *out))
        return out

    def cfl_dt(self, h, courant=0.4):
        return lib().fl_model_stream_cfl_dt(self.handle, h, courant)

    def close(self):
        if self.handle:
            lib().fl_model_stream_close(self.handle)
            self.handle = None


def cfl_dt(model, h, courant=0.4):
    return lib().fl_cfl_dt(model[0].size, model[0], model[2], h, courant)


def device_name():
    buf = C.create_string_buffer(256)
    _check(lib().fl_device_name(buf, 256))
    return buf.value.decode()


# ------------------------------------------------------------------ driver.h

class Driver:
    """The reference back-end API on the GPU, plus the fl_hip.h extensions.

    Mirrors original/driver.h:8-22 call for call; errors raise FletcherError
    (the C layer records them instead of exiting)."""

    def __init__(self, variant="fused", precision="fp32", sync=True, tiles=None):
        self.L = lib()
        _check(self.L.fl_set_variant(VARIANT[variant]))
        _check(self.L.fl_set_precision(PRECISION[precision]))
        _check(self.L.fl_set_sync(1 if sync else 0))
        if tiles is not None:
            _check(self.L.fl_set_tiles(*tiles))
        self.shape = None

    def _status(self):
        if self.L.fl_status() != 0:
            msg = self.L.fl_last_error().decode()
            self.L.fl_clear_error()
            raise FletcherError(msg)

    def initialize(self, shape, d, dt, model, fields=None, bord=BORD):
        sz, sy, sx = shape
        self.shape, self.bord, self.d, self.dt = tuple(shape), bord, tuple(d), dt
        model = [_c32(m) for m in model]
        if fields is None:
            ptrs = [None] * 4
        else:
            fields = [_c32(f) for f in fields]
            ptrs = [f.ctypes.data for f in fields]
        self.L.fl_clear_error()
        self.L.DRIVER_Initialize(sx, sy, sz, bord, d[0], d[1], d[2], dt, *model, *ptrs)
        self._status()

    def initialize_streamed(self, stream, d, dt, bord=BORD):
        """DRIVER_Initialize from a ModelStream: the model goes to the device in z slabs and the
        coefficient records are computed there; no whole model arrays on the host."""
        sz, sy, sx = stream.shape
        self.shape, self.bord, self.d, self.dt = tuple(stream.shape), bord, tuple(d), dt
        fill = C.cast(self.L.fl_model_stream_fill, C.c_void_p)
        self.L.fl_clear_error()
        _check(self.L.fl_initialize_streamed(sx, sy, sz, bord, d[0], d[1], d[2], dt, fill, stream.handle))

    def set_device_precompute(self, on):
        """DRIVER_Initialize computes the coefficient records on the device (1) or the host (0)."""
        _check(self.L.fl_set_device_precompute(1 if on else 0))

    def insert_source(self, it, isource, src):
        self.L.DRIVER_InsertSource(self.dt, it, isource, None, None, src)
        self._status()

    def propagate(self, it):
        sz, sy, sx = self.shape
        self.L.DRIVER_Propagate(sx, sy, sz, self.bord, self.d[0], self.d[1], self.d[2], self.dt, it,
                                None, None, None, None)
        self._status()

    def update_pointers(self, out=None):
        sz, sy, sx = self.shape
        out = np.empty(self.shape, np.float32) if out is None else out
        self.L.DRIVER_Update_pointers(sx, sy, sz, out.ctypes.data)
        self._status()
        return out

    def finalize(self):
        self.L.DRIVER_Finalize()

    def field(self, which):
        out = np.empty(self.shape, np.float32)
        _check(self.L.fl_get_field(FIELD[which], out))
        return out

    def set_field(self, which, arr):
        _check(self.L.fl_set_field(FIELD[which], _c32(arr)))

    def record_traces(self, iz, every, max_traces):
        _check(self.L.fl_record_traces(iz, every, max_traces))
        self._rec = (every, max_traces)

    def traces(self):
        sz, sy, sx = self.shape
        every, m = self._rec
        out = np.empty((m, sy - 2 * self.bord, sx - 2 * self.bord), np.float32)
        n = _check(self.L.fl_get_traces(out, m))
        return out[:n]

    def run(self, st, isource, it0=0, graph_chunk=50, src=None):
        """Steps it0+1..it0+st of the reference loop (insert Source(dt,it-1), propagate)."""
        if src is None:
            src = sources(self.dt, st, it0)
        _check(self.L.fl_run_steps(it0, st, isource, _c32(src), graph_chunk))

    def reset_fields(self):
        """All fields and recorded traces back to zero, device state and graphs kept."""
        _check(self.L.fl_reset_fields())

    def step_loop(self, st, isource, it0=0):
        """The same steps issued call by call through DRIVER_InsertSource/DRIVER_Propagate."""
        for it in range(it0 + 1, it0 + st + 1):
            self.insert_source(it - 1, isource, source(self.dt, it - 1))
            self.propagate(it)

    def records(self):
        """The device coefficient records (raw bytes, one record per pitched cell)."""
        out = np.empty(_check(self.L.fl_record_bytes()), np.uint8)
        _check(self.L.fl_get_records(out.ctypes.data))
        return out

    def derivative_volume(self, k):
        """Two-pass first-derivative volume k (0 pDx, 1 pDy, 2 qDx, 3 qDy) of the last step."""
        out = np.empty(self.shape, np.float32)
        _check(self.L.fl_get_derivative_volume(k, out))
        return out

    def set_variant(self, v):
        _check(self.L.fl_set_variant(VARIANT[v]))

    def set_sync(self, on):
        _check(self.L.fl_set_sync(1 if on else 0))

    def set_tiles(self, ty, zt):
        _check(self.L.fl_set_tiles(ty, zt))

    def tiles(self):
        a, b = C.c_int(), C.c_int()
        self.L.fl_get_tiles(C.byref(a), C.byref(b))
        return a.value, b.value

    def set_stream_hints(self, on):
        _check(self.L.fl_set_stream_hints(1 if on else 0))

    def stream_hints(self):
        return bool(self.L.fl_get_stream_hints())

    def set_tile_strip(self, sw):
        """Tile columns per strip of the step's workgroup order (0: the grid's default)."""
        _check(self.L.fl_set_tile_strip(sw))

    def tile_strip(self):
        return self.L.fl_get_tile_strip()

    def set_lane_pairs(self, on):
        """Both fields of a cell per lane in the 16-bit recompute forms' step (same results)."""
        _check(self.L.fl_set_lane_pairs(1 if on else 0))

    def lane_pairs(self):
        return bool(self.L.fl_get_lane_pairs())

    def autotune(self, steps=5, cache=None):
        _check(self.L.fl_autotune(steps, cache.encode() if cache else None))
        return self.tiles()

    def set_timing(self, on):
        _check(self.L.fl_set_timing(1 if on else 0))

    def kernel_times(self):
        ms, n = C.c_double(), C.c_int()
        _check(self.L.fl_kernel_times(C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def region_begin(self):
        _check(self.L.fl_region_begin())

    def region_end(self):
        ms = C.c_double()
        _check(self.L.fl_region_end(C.byref(ms)))
        return ms.value

    def synchronize(self):
        _check(self.L.fl_synchronize())

    def bytes_per_cell(self):
        return self.L.fl_bytes_per_cell()

    def device_bytes(self):
        return self.L.fl_device_bytes()


# ------------------------------------------------------------------ batched shots

class _Shot(C.Structure):
    _fields_ = [("sx", C.c_int), ("sy", C.c_int), ("sz", C.c_int), ("isource", C.c_int),
                ("rec_iz", C.c_int)] + [(n, C.c_void_p) for n in MODEL_NAMES]


class Batch:
    """Shots on their own sub-grids, advanced together (fl_batch_*)."""

    def __init__(self, shots, d, dt, precision="fp32", bord=BORD, tiles=None):
        """shots: list of dicts with 'model' (6 arrays), 'isource', 'rec_iz' (-1 none)."""
        self.L = lib()
        self.bord, self.dt = bord, dt
        self._keep = []
        arr = (_Shot * len(shots))()
        self.shapes = []
        for k, s in enumerate(shots):
            m = [_c32(a) for a in s["model"]]
            self._keep.append(m)
            sz, sy, sx = m[0].shape
            self.shapes.append((sz, sy, sx))
            arr[k].sx, arr[k].sy, arr[k].sz = sx, sy, sz
            arr[k].isource = s["isource"]
            arr[k].rec_iz = s.get("rec_iz", -1)
            for n, a in zip(MODEL_NAMES, m):
                setattr(arr[k], n, a.ctypes.data)
        if tiles is not None:
            _check(self.L.fl_batch_set_tiles(*tiles))
        _check(self.L.fl_batch_create(len(shots), C.cast(arr, C.c_void_p), bord, d[0], d[1], d[2], dt,
                                      PRECISION[precision]))
        self._keep = None  # the device holds its own copies now
        self.rec = (0, 0)

    def run(self, st, it0=0, graph_chunk=50, rec_every=0, max_traces=0, src=None):
        if src is None:
            src = sources(self.dt, st, it0)
        _check(self.L.fl_batch_run(it0, st, _c32(src), graph_chunk, rec_every, max_traces))
        if rec_every:
            self.rec = (rec_every, max_traces)

    def reset(self):
        """Every shot's fields and traces back to zero."""
        _check(self.L.fl_batch_reset())

    def field(self, shot, which):
        out = np.empty(self.shapes[shot], np.float32)
        _check(self.L.fl_batch_get_field(shot, FIELD[which], out))
        return out

    def traces(self, shot):
        sz, sy, sx = self.shapes[shot]
        m = self.rec[1]
        out = np.empty((m, sy - 2 * self.bord, sx - 2 * self.bord), np.float32)
        n = _check(self.L.fl_batch_get_traces(shot, out, m))
        return out[:n]

    def autotune(self, steps=4, cache=None):
        _check(self.L.fl_batch_autotune(steps, cache.encode() if cache else None))

    def set_tiles(self, ty, zt):
        _check(self.L.fl_batch_set_tiles(ty, zt))

    def set_persistent(self, wgs):
        """0: one workgroup per item; 1: the persistent form over the resident workgroups;
        > 1: the persistent form over at most wgs workgroups."""
        _check(self.L.fl_batch_set_persistent(wgs))

    def set_lane_pairs(self, on):
        """Both fields of a cell per lane in the 16-bit recompute forms' step (same results;
        the setting is the library's, shared with Driver.set_lane_pairs)."""
        _check(self.L.fl_set_lane_pairs(1 if on else 0))

    def tiles(self):
        ty, zt = C.c_int(), C.c_int()
        _check(self.L.fl_batch_get_tiles(C.byref(ty), C.byref(zt)))
        return ty.value, zt.value

    def last_ms(self):
        ms = C.c_double()
        self.L.fl_batch_kernel_ms(C.byref(ms))
        return ms.value

    def destroy(self):
        self.L.fl_batch_destroy()


class Ref64:
    """FP64 accuracy reference on the device (fl_ref64_*): the der1der1 time loop in double
    with exact weights and FP64 coefficients, reference layout.  One instance at a time."""

    def __init__(self, shape, d, dt, model, bord=BORD):
        self.L = lib()
        self.shape, self.bord, self.dt = tuple(shape), bord, dt
        m = [_c32(a) for a in model]
        sz, sy, sx = self.shape
        _check(self.L.fl_ref64_create(sx, sy, sz, bord, d[0], d[1], d[2], dt, *m))
        self.slots = 0

    def run(self, st, isource, it0=0, src=None, rec_iz=-1, rec_every=0, max_traces=0):
        if src is None:
            src = sources(self.dt, st, it0)
        _check(self.L.fl_ref64_run(it0, st, isource, _c32(src), rec_iz, rec_every, max_traces))
        if rec_every:
            self.slots = max_traces

    def field(self, which):
        out = np.empty(self.shape, np.float64)
        _check(self.L.fl_ref64_get_field(FIELD[which], out))
        return out

    def traces(self):
        sz, sy, sx = self.shape
        out = np.empty((self.slots, sy - 2 * self.bord, sx - 2 * self.bord), np.float64)
        n = _check(self.L.fl_ref64_get_traces(out, self.slots))
        return out[:n]

    def destroy(self):
        self.L.fl_ref64_destroy()
