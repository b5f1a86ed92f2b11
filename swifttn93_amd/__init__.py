"""swifttn93_amd: all-pairs TN93 distances (tn93 v1.1.0 standard mode) on one MI355X.

The product is the native library swifttn93_amd/lib/libtn93g.so (C ABI in include/tn93g.h) and the tn93-compatible
CLI swifttn93_amd/bin/tn93g.  This module is a thin ctypes binding for tests and benchmarks; it has no compute
fallback: without the built library every call raises.
"""
import ctypes
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "lib" / "libtn93g.so"
CLI_PATH = PKG / "bin" / "tn93g"

RESOLVE, AVERAGE, SKIP, GAPMM, SUBSET = 0, 1, 2, 3, 4
TRIANGLE, CROSS, FST, APPEND = 0, 1, 2, 3
BACKEND_AUTO, BACKEND_POPC, BACKEND_MFMA = 0, 1, 2
BACKEND_GEMM = 3  # reported in backend_used only: the counts of a long dense-N alignment as rocBLAS int8 GEMMs
FLAG_EXACT, FLAG_ALL_SLOW, FLAG_AUTOTUNE = 1, 2, 4

MODES = {"resolve": RESOLVE, "average": AVERAGE, "skip": SKIP, "gapmm": GAPMM}


class Params(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int32), ("kind", ctypes.c_int32), ("n1", ctypes.c_int64),
                ("subset_mask", ctypes.c_uint32), ("backend", ctypes.c_int32),
                ("resolve_fraction", ctypes.c_double), ("min_overlap", ctypes.c_int64),
                ("threshold", ctypes.c_double), ("min_distance", ctypes.c_double), ("flags", ctypes.c_uint32),
                ("n2", ctypes.c_int64), ("tile_rows", ctypes.c_int32)]


class Result(ctypes.Structure):
    _fields_ = [("exact", ctypes.c_int32), ("nrows", ctypes.c_int32), ("hist", (ctypes.c_uint64 * 200) * 3),
                ("n_links", ctypes.c_int64), ("link_row", ctypes.POINTER(ctypes.c_int64)),
                ("link_col", ctypes.POINTER(ctypes.c_int64)), ("link_dist", ctypes.POINTER(ctypes.c_double)),
                ("link_dist_lo", ctypes.POINTER(ctypes.c_double)), ("link_dist_hi", ctypes.POINTER(ctypes.c_double)),
                ("mean_lo", ctypes.c_double * 3), ("mean_hi", ctypes.c_double * 3), ("weighted", ctypes.c_double * 3),
                ("max_lo", ctypes.c_double * 3), ("max_hi", ctypes.c_double * 3), ("found_links", ctypes.c_int64),
                ("skipped", ctypes.c_int64), ("n_replayed", ctypes.c_int64), ("n_uncertain", ctypes.c_int64),
                ("n_host_exact", ctypes.c_int64), ("gpu_ms", ctypes.c_double), ("pair_kernel_ms", ctypes.c_double),
                ("tail_kernel_ms", ctypes.c_double), ("backend_used", ctypes.c_int32),
                ("tile_rows_used", ctypes.c_int32), ("folded_columns", ctypes.c_int64)]


_LINK_PTR_OFFSET = Result.link_row.offset  # link_row .. link_dist_hi: five consecutive pointers
assert all(getattr(Result, f).offset == _LINK_PTR_OFFSET + 8 * k
           for k, f in enumerate(["link_row", "link_col", "link_dist", "link_dist_lo", "link_dist_hi"]))

EXPORTS = ["tn93g_open", "tn93g_close", "tn93g_error", "tn93g_set_upload", "tn93g_set_free", "tn93g_run",
           "tn93g_result_free", "tn93g_version", "tn93g_run_batch", "tn93g_set_create", "tn93g_set_append",
           "tn93g_set_size", "tn93g_run_bytes", "tn93g_result_certify", "tn93g_run_noalloc"]
NOALLOC_SCRATCH_BYTES = 16 << 20  # TN93G_NOALLOC_SCRATCH_BYTES

_lib = None


def load():
    """Loads libtn93g.so (raises if it has not been built: there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python3 -m swifttn93_amd.build`")
    lib = ctypes.CDLL(str(LIB_PATH))
    lib.tn93g_open.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
    lib.tn93g_close.argtypes = [ctypes.c_void_p]
    lib.tn93g_error.argtypes = [ctypes.c_void_p]
    lib.tn93g_error.restype = ctypes.c_char_p
    lib.tn93g_set_upload.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                     ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
    lib.tn93g_set_free.argtypes = [ctypes.c_void_p]
    lib.tn93g_set_create.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(ctypes.c_void_p)]
    lib.tn93g_set_append.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
    lib.tn93g_set_size.argtypes = [ctypes.c_void_p]
    lib.tn93g_set_size.restype = ctypes.c_int64
    lib.tn93g_run_bytes.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(Params),
                                    ctypes.POINTER(ctypes.c_int64)]
    lib.tn93g_run.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(Params),
                              ctypes.POINTER(ctypes.POINTER(Result))]
    lib.tn93g_result_free.argtypes = [ctypes.POINTER(Result)]
    lib.tn93g_run_batch.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p),
                                    ctypes.POINTER(Params), ctypes.POINTER(ctypes.POINTER(Result))]
    lib.tn93g_result_certify.argtypes = [ctypes.POINTER(Result), ctypes.c_int]
    lib.tn93g_run_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(Params), ctypes.POINTER(Result),
                                      ctypes.c_int64]
    lib.tn93g_version.restype = ctypes.c_char_p
    _lib = lib
    return lib


def make_params(mode=RESOLVE, kind=TRIANGLE, n1=0, subset_mask=0, backend=BACKEND_AUTO, resolve_fraction=1.0,
                min_overlap=100, threshold=0.015, min_distance=-1e-10, flags=0, n2=0, tile_rows=0):
    return Params(mode, kind, n1, subset_mask, backend, resolve_fraction, min_overlap, threshold, min_distance, flags,
                  n2, tile_rows)


class Engine:
    """One GPU context (tn93g_open)."""

    def __init__(self, device=0):
        self.lib = load()
        self.ctx = ctypes.c_void_p()
        rc = self.lib.tn93g_open(device, ctypes.byref(self.ctx))
        if rc:
            raise RuntimeError(f"tn93g_open failed ({rc}): no usable gfx950 device")

    def _check(self, rc):
        if rc:
            raise RuntimeError(f"tn93g error {rc}: {self.lib.tn93g_error(self.ctx).decode()}")

    def upload(self, codes, weights=None):
        codes = np.ascontiguousarray(codes, dtype=np.uint8)
        n, L = codes.shape
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.int64)
        s = ctypes.c_void_p()
        self._check(self.lib.tn93g_set_upload(self.ctx, codes.ctypes.data, n, L,
                                              None if w is None else w.ctypes.data, ctypes.byref(s)))
        return SeqSet(self, s, n, L)

    def create(self, L, capacity):
        """An empty set of sequences of L sites with room for `capacity` sequences (tn93g_set_create)."""
        s = ctypes.c_void_p()
        self._check(self.lib.tn93g_set_create(self.ctx, L, capacity, ctypes.byref(s)))
        return SeqSet(self, s, 0, L)

    def run_bytes(self, seqset, **kw):
        """Device bytes a run of these parameters works in (tn93g_run_bytes)."""
        p = make_params(**kw)
        b = ctypes.c_int64()
        self._check(self.lib.tn93g_run_bytes(self.ctx, seqset.handle, ctypes.byref(p), ctypes.byref(b)))
        return b.value

    def run(self, seqset, mode=RESOLVE, kind=TRIANGLE, n1=0, subset_mask=0, backend=BACKEND_AUTO,
            resolve_fraction=1.0, min_overlap=100, threshold=0.015, min_distance=-1e-10, flags=0, n2=0,
            certify=False, tile_rows=0):
        """One tn93g_run.  With certify, a result whose link distances do not print as tn93's would (%g at both
        ends of their enclosures, tn93g_result_certify) is replaced by the run again with FLAG_EXACT, so every
        returned link_dist prints as tn93 prints it."""
        p = make_params(mode, kind, n1, subset_mask, backend, resolve_fraction, min_overlap, threshold, min_distance,
                        flags, n2, tile_rows)
        r = ctypes.POINTER(Result)()
        self._check(self.lib.tn93g_run(self.ctx, seqset.handle, ctypes.byref(p), ctypes.byref(r)))
        ok = None
        if certify:
            ok = self.lib.tn93g_result_certify(r, 6) == 1
            if not ok:
                self.lib.tn93g_result_free(r)
                p.flags |= FLAG_EXACT
                r = ctypes.POINTER(Result)()
                self._check(self.lib.tn93g_run(self.ctx, seqset.handle, ctypes.byref(p), ctypes.byref(r)))
                ok = self.lib.tn93g_result_certify(r, 6) == 1
        d = _to_py(r, _ResultOwner(self.lib, r))
        d["certified"] = ok  # None unless certify was asked for
        return d

    def run_noalloc(self, seqset, link_capacity=1 << 16, **kw):
        """tn93g_run_noalloc: the run in the device's fixed scratch, its links written into arrays allocated here
        beforehand (the library allocates nothing).  Raises with the library's message when the run does not fit."""
        p = make_params(**kw)
        arrs = [np.zeros(max(link_capacity, 1), dtype=t) for t in (np.int64, np.int64, np.float64, np.float64,
                                                                   np.float64)]
        r = Result()
        r.link_row = arrs[0].ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
        r.link_col = arrs[1].ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
        r.link_dist = arrs[2].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        r.link_dist_lo = arrs[3].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        r.link_dist_hi = arrs[4].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        self._check(self.lib.tn93g_run_noalloc(self.ctx, seqset.handle, ctypes.byref(p), ctypes.byref(r),
                                               link_capacity))
        n = r.n_links
        return {
            "exact": bool(r.exact), "certified": self.lib.tn93g_result_certify(ctypes.byref(r), 6) == 1,
            "hist": np.frombuffer(r.hist, dtype=np.uint64).reshape(3, 200).copy(),
            "link_row": arrs[0][:n], "link_col": arrs[1][:n], "link_dist": arrs[2][:n], "link_lo": arrs[3][:n],
            "link_hi": arrs[4][:n],
            "mean_lo": list(r.mean_lo), "mean_hi": list(r.mean_hi), "weighted": list(r.weighted),
            "max_lo": list(r.max_lo), "max_hi": list(r.max_hi), "found_links": r.found_links, "skipped": r.skipped,
            "n_replayed": r.n_replayed, "n_uncertain": r.n_uncertain, "n_host_exact": r.n_host_exact,
            "gpu_ms": r.gpu_ms, "pair_kernel_ms": r.pair_kernel_ms, "tail_kernel_ms": r.tail_kernel_ms,
            "folded_columns": r.folded_columns,
        }

    def run_batch(self, jobs):
        """jobs: list of (SeqSet, dict of run() keyword arguments).  One tn93g_run_batch call; a list of results."""
        n = len(jobs)
        sets = (ctypes.c_void_p * max(n, 1))(*[s.handle for s, _ in jobs])
        params = (Params * max(n, 1))(*[make_params(**kw) for _, kw in jobs])
        outs = (ctypes.POINTER(Result) * max(n, 1))()
        self._check(self.lib.tn93g_run_batch(self.ctx, n, sets, params, outs))
        owners = [_ResultOwner(self.lib, outs[k]) for k in range(n)]
        return [_to_py(outs[k], owners[k]) for k in range(n)]

    def close(self):
        if self.ctx:
            self.lib.tn93g_close(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SeqSet:
    def __init__(self, eng, handle, n, L):
        self.eng, self.handle, self.n, self.L = eng, handle, n, L

    def append(self, codes, weights=None):
        """Appends sequences (tn93g_set_append); returns the index of the first one."""
        codes = np.ascontiguousarray(codes, dtype=np.uint8)
        n, L = codes.shape
        assert L == self.L
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.int64)
        lib = self.eng.lib
        first = lib.tn93g_set_size(self.handle)
        rc = lib.tn93g_set_append(self.handle, codes.ctypes.data, n, None if w is None else w.ctypes.data)
        if rc:
            raise RuntimeError(f"tn93g_set_append failed ({rc}): {lib.tn93g_error(self.eng.ctx).decode()}")
        self.n = lib.tn93g_set_size(self.handle)
        return first

    def size(self):
        return self.eng.lib.tn93g_set_size(self.handle)

    def free(self):
        if self.handle:
            self.eng.lib.tn93g_set_free(self.handle)
            self.handle = None


class _ResultOwner:
    """Frees a tn93g_result when the last array viewing its link block is gone."""

    def __init__(self, lib, ptr):
        self.lib, self.ptr = lib, ptr

    def __del__(self):
        if self.ptr:
            try:
                self.lib.tn93g_result_free(self.ptr)
            except Exception:  # interpreter shutdown
                pass
            self.ptr = None


class _View:
    def __init__(self, owner, addr, n, typestr):
        self.owner = owner  # keeps the result alive as long as the array
        self.__array_interface__ = {"shape": (n,), "typestr": typestr, "data": (addr, False), "version": 3}


def _arr(owner, ptr, n, dt):
    """The n values at ptr as a numpy array that shares the library's memory (no copy)."""
    if n == 0:
        return np.zeros(0, dtype=dt)
    return np.asarray(_View(owner, ctypes.cast(ptr, ctypes.c_void_p).value, n, np.dtype(dt).str))


_SMALL_LINKS = 4096  # up to this many links are copied out (cheaper than views), after which the result is freed


def _copy(ptr, n, ctype, dt):
    """A copy of the n values at ptr."""
    if n == 0:
        return np.zeros(0, dtype=dt)
    return np.frombuffer((ctype * n).from_address(ctypes.cast(ptr, ctypes.c_void_p).value), dtype=dt).copy()


def _link_views(owner, r, n):
    """The five link arrays as views of the result's link block: one array over the block, sliced, when the arrays lie
    in the library's layout (five arrays of one capacity, one after the other), else one view per array."""
    a = (ctypes.c_size_t * 5).from_buffer(r, _LINK_PTR_OFFSET)  # the five pointers as integers
    cap8 = a[1] - a[0]
    if cap8 >= 8 * n and all(a[k] - a[0] == k * cap8 for k in range(5)):
        cap = cap8 // 8
        blk = np.asarray(_View(owner, a[0], 5 * cap, "<i8"))
        f = blk.view(np.float64)
        return [blk[:n], blk[cap:cap + n], f[2 * cap:2 * cap + n], f[3 * cap:3 * cap + n], f[4 * cap:4 * cap + n]]
    return [_arr(owner, r.link_row, n, np.int64), _arr(owner, r.link_col, n, np.int64),
            _arr(owner, r.link_dist, n, np.float64), _arr(owner, r.link_dist_lo, n, np.float64),
            _arr(owner, r.link_dist_hi, n, np.float64)]


def _to_py(rp, owner):
    r = rp.contents
    n = r.n_links
    if n <= _SMALL_LINKS:
        lk = [_copy(r.link_row, n, ctypes.c_int64, np.int64), _copy(r.link_col, n, ctypes.c_int64, np.int64),
              _copy(r.link_dist, n, ctypes.c_double, np.float64), _copy(r.link_dist_lo, n, ctypes.c_double, np.float64),
              _copy(r.link_dist_hi, n, ctypes.c_double, np.float64)]
    else:
        lk = _link_views(owner, r, n)
    return {
        "exact": bool(r.exact),
        "hist": np.frombuffer(r.hist, dtype=np.uint64).reshape(3, 200).copy(),
        "link_row": lk[0], "link_col": lk[1], "link_dist": lk[2], "link_lo": lk[3], "link_hi": lk[4],
        "mean_lo": list(r.mean_lo), "mean_hi": list(r.mean_hi), "weighted": list(r.weighted),
        "max_lo": list(r.max_lo), "max_hi": list(r.max_hi), "found_links": r.found_links, "skipped": r.skipped,
        "n_replayed": r.n_replayed, "n_uncertain": r.n_uncertain, "n_host_exact": r.n_host_exact,
        "gpu_ms": r.gpu_ms, "pair_kernel_ms": r.pair_kernel_ms, "tail_kernel_ms": r.tail_kernel_ms,
        "backend_used": r.backend_used, "tile_rows_used": r.tile_rows_used, "folded_columns": r.folded_columns,
    }
