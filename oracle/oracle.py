"""ORACLE — test infrastructure only (ctypes wrapper around liboracle.so and the
reference build oracle/_ref/libref_<variant>.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, and only as the checker or as the CPU baseline; the product never does.

liboracle.so restates the reference's OpenMP propagate path (see fl_oracle.c for
the file:line map).  libref_*.so are the unmodified reference sources compiled by
oracle/Makefile plus ref_harness.c.
"""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")
VARIANTS = ("original", "der1der1", "der1der1hm", "der1der1lm")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_lib = None


def build(ref=True):
    """Compile liboracle.so (and the reference libraries when the sources exist)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if ref:
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build(ref=False)
        L = C.CDLL(path)
        L.oracle_source.restype = C.c_float
        L.oracle_source.argtypes = [C.c_float, C.c_int]
        for sfx, ptr in (("f32", _f32p), ("f64", _f64p)):
            f = getattr(L, "oracle_run_" + sfx)
            real = C.c_float if sfx == "f32" else C.c_double
            f.restype = C.c_int
            f.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, real, real, real, real,
                          C.c_int, C.c_long, _f32p, C.POINTER(C.c_void_p), ptr, ptr, ptr, ptr,
                          C.c_int, C.c_int, C.c_void_p]
            for name in ("der1", "der2"):
                g = getattr(L, "oracle_%s_%s" % (name, sfx))
                g.restype = real
                g.argtypes = [ptr, C.c_long, C.c_long, real]
            g = getattr(L, "oracle_dercross_" + sfx)
            g.restype = real
            g.argtypes = [ptr, C.c_long, C.c_long, C.c_long, real]
            p = getattr(L, "oracle_precompute_" + sfx)
            p.restype = None
            p.argtypes = [C.c_long] + [_f32p] * 6 + [C.POINTER(C.c_void_p)]
        L.oracle_set_threads.argtypes = [C.c_int]
        L.oracle_get_threads.restype = C.c_int
        _lib = L
    return _lib


def set_threads(n):
    lib().oracle_set_threads(int(n))


def get_threads():
    return lib().oracle_get_threads()


def source(dt, it):
    return lib().oracle_source(C.c_float(dt), int(it))


def sources(dt, st):
    return np.array([source(dt, k) for k in range(st)], dtype=np.float32)


def _ptrs(arrs):
    return (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


def precompute(model, fp="f32"):
    """10 coefficient volumes (precomp order) from the 6 model arrays
    (vpz, vsv, epsilon, delta, phi, theta), each float32 C-contiguous."""
    dt = np.float32 if fp == "f32" else np.float64
    n = model[0].size
    out = [np.empty(model[0].shape, dt) for _ in range(10)]
    getattr(lib(), "oracle_precompute_" + fp)(n, *[np.ascontiguousarray(m, np.float32) for m in model],
                                             _ptrs(out))
    return out


def run(variant, shape, bord, d, dt, st, isource, coefs, fields=None, fp="f32",
        rec_iz=-1, rec_every=0, src=None):
    """Run st steps of the reference time loop on the oracle.

    variant: "original" (direct DerCross), "der1der1" (two-pass first-of-first), or its
    loop orders "der1der1hm" (first derivatives formed plane by plane in z order) and
    "der1der1lm" (first derivatives in a 9-plane ring).
    shape: (sz, sy, sx); d: (dx, dy, dz); coefs: 10 arrays of dtype fp.
    Returns dict with pp, pc, qp, qc (pc/qc = field at t=st*dt) and traces.
    """
    sz, sy, sx = shape
    dtype = np.float32 if fp == "f32" else np.float64
    if fields is None:
        fields = [np.zeros(shape, dtype) for _ in range(4)]
    else:
        fields = [np.array(f, dtype=dtype, copy=True) for f in fields]
    if src is None:
        src = sources(dt, st)
    src = np.ascontiguousarray(src, np.float32)
    coefs = [np.ascontiguousarray(c, dtype) for c in coefs]
    traces = None
    tptr = None
    if rec_every > 0:
        nt = st // rec_every
        traces = np.zeros((nt, sy - 2 * bord, sx - 2 * bord), dtype)
        tptr = traces.ctypes.data
    f = getattr(lib(), "oracle_run_" + fp)
    v = {"original": 0, "der1der1": 1, "der1der1hm": 2, "der1der1lm": 3}[variant]
    ntr = f(v, sx, sy, sz, bord, d[0], d[1], d[2], dt, st, isource, src, _ptrs(coefs),
            fields[0], fields[1], fields[2], fields[3], rec_iz, rec_every, tptr)
    if ntr < 0:
        raise MemoryError("oracle scratch allocation failed")
    return {"pp": fields[0], "pc": fields[1], "qp": fields[2], "qc": fields[3], "traces": traces}


def rel_l2(a, b):
    """Relative L2 error of a against b, accumulated in float64."""
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    den = np.sqrt(np.dot(b, b))
    num = np.sqrt(np.dot(a - b, a - b))
    return float(num / den) if den > 0 else float(num)


# ---------------------------------------------------------------- reference build


def ref_available(variant="original"):
    return os.path.exists(os.path.join(REF_DIR, "libref_%s.so" % variant))


class Ref:
    """The reference C/OpenMP program of one variant, loaded from oracle/_ref."""

    def __init__(self, variant="original"):
        path = os.path.join(REF_DIR, "libref_%s.so" % variant)
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        # RTLD_LOCAL (the default) keeps each variant's identical symbol names apart
        L = C.CDLL(path)
        L.ref_main.argtypes = [C.c_int, C.POINTER(C.c_char_p)]
        L.ref_main.restype = C.c_int
        L.ref_capture_size.argtypes = [C.POINTER(C.c_int), C.POINTER(C.c_float)]
        L.ref_capture_size.restype = C.c_long
        L.ref_capture_get.argtypes = [C.c_int, _f32p]
        L.ref_setup.argtypes = [C.c_int] * 4 + [C.c_float] * 4 + [_f32p] * 6
        L.ref_steps.argtypes = [C.c_int, C.c_int, C.c_int]
        L.ref_steps.restype = C.c_double
        L.ref_fetch.argtypes = [C.c_int, _f32p]
        L.ref_fetch_coef.argtypes = [C.c_int, _f32p]
        L.ref_source.argtypes = [C.c_float, C.c_int]
        L.ref_source.restype = C.c_float
        L.ref_random_boundary.argtypes = [C.c_uint] + [C.c_int] * 8 + [_f32p, _f32p]
        L.ref_set_threads.argtypes = [C.c_int]
        L.ref_set_run_model.argtypes = [C.c_int]
        self.L = L
        self.variant = variant
        self.shape = None

    def main(self, args, run_model=True, workdir=None):
        """Run the reference main() with argv = ["ref"] + args inside workdir (the reference
        writes Report.csv and, for der1der1, TTI.rsf there).  Returns captures dict."""
        self.L.ref_set_run_model(1 if run_model else 0)
        self.L.ref_reset_rand()
        argv = [b"ModelagemFletcher.exe"] + [str(a).encode() for a in args]
        arr = (C.c_char_p * len(argv))(*argv)
        cwd = os.getcwd()
        tmp = workdir or tempfile.mkdtemp(prefix="fl_ref_")
        os.chdir(tmp)
        try:
            self.L.ref_main(len(argv), arr)
        finally:
            os.chdir(cwd)
        dims = (C.c_int * 5)()
        dtv = C.c_float()
        n = self.L.ref_capture_size(dims, C.byref(dtv))
        sx, sy, sz, st, isrc = list(dims)
        shape = (sz, sy, sx)
        out = {"shape": shape, "st": st, "isource": isrc, "dt": dtv.value, "workdir": tmp}
        names = ["vpz", "vsv", "epsilon", "delta", "phi", "theta"]
        for k, nm in enumerate(names):
            a = np.empty(n, np.float32)
            self.L.ref_capture_get(k, a)
            out[nm] = a.reshape(shape)
        if run_model:
            out["coefs"] = []
            for k in range(10):
                a = np.empty(n, np.float32)
                self.L.ref_capture_get(10 + k, a)
                out["coefs"].append(a.reshape(shape))
            for k, nm in enumerate(["pp", "pc", "qp", "qc"]):
                a = np.empty(n, np.float32)
                self.L.ref_capture_get(20 + k, a)
                out[nm] = a.reshape(shape)
        return out

    def setup(self, shape, bord, d, dt, model):
        sz, sy, sx = shape
        self.shape = shape
        model = [np.ascontiguousarray(m, np.float32) for m in model]
        rc = self.L.ref_setup(sx, sy, sz, bord, d[0], d[1], d[2], dt, *model)
        if rc != 0:
            raise MemoryError("reference setup failed")

    def steps(self, it0, n, isource):
        """Run steps it0+1..it0+n; returns Propagate+Swap wall seconds (model.c:82-88)."""
        return self.L.ref_steps(it0, n, isource)

    def fetch(self, which):
        a = np.empty(int(np.prod(self.shape)), np.float32)
        self.L.ref_fetch({"pp": 0, "pc": 1, "qp": 2, "qc": 3}[which], a)
        return a.reshape(self.shape)

    def fetch_coefs(self):
        out = []
        for k in range(10):
            a = np.empty(int(np.prod(self.shape)), np.float32)
            self.L.ref_fetch_coef(k, a)
            out.append(a.reshape(self.shape))
        return out

    def source(self, dt, it):
        return self.L.ref_source(dt, it)

    def random_boundary(self, seed, shape, n3, bord, absorb, vpz, vsv):
        sz, sy, sx = shape
        vpz = np.ascontiguousarray(vpz, np.float32).copy()
        vsv = np.ascontiguousarray(vsv, np.float32).copy()
        self.L.ref_random_boundary(seed, sx, sy, sz, n3[0], n3[1], n3[2], bord, absorb, vpz, vsv)
        return vpz, vsv

    def set_threads(self, n):
        self.L.ref_set_threads(int(n))
