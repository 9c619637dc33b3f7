"""Python host binding of libgcfg.so (the MI355X binary-MRF min-cut library).

Mirrors the reference's host API for this path: ``Graph`` follows the
``GC_*`` functions of ``GraphCut.cu:187-568`` (create = GC_Init allocation,
``set_dataterms``/``set_edges``/``set_penalty``/``build``/``optimize``/``close``),
``DataTerm`` follows ``DataTerm.h:9-19``. ``Batch`` and ``Server`` expose the
batched and serving forms. Device memory is plain HIP memory managed through
``libamdhip64`` (``DeviceBuffer``); there is no CPU fallback: every solve runs
the HIP kernels of libgcfg.so and fails loudly if the library or a GPU is
missing.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# GCFG_LIBRARY selects another build of the library (same ABI).
LIB_PATH = os.environ.get("GCFG_LIBRARY") or os.path.join(_HERE, "libgcfg.so")

GCFG_OK = 0
GCFG_ERR_INVALID = -1
GCFG_ERR_HIP = -2
GCFG_ERR_NOMEM = -3
GCFG_ERR_STATE = -4
GCFG_ERR_NOT_CONVERGED = -5
GCFG_ERR_BUDGET = -6
SEED_NONE, SEED_FG, SEED_BG = 0, 1, 2
DT_MODEL_HISTOGRAM, DT_MODEL_KDE = 0, 1
LABELS_INT32, LABELS_BITS = 0, 1

_I32P = ctypes.POINTER(ctypes.c_int32)
_I64P = ctypes.POINTER(ctypes.c_int64)
_VP = ctypes.c_void_p


class GcfgError(RuntimeError):
    def __init__(self, rc: int, what: str, msg: str):
        super().__init__(f"{what} failed ({rc}): {msg}")
        self.rc = rc


class Options(ctypes.Structure):
    _fields_ = [("sweeps_per_gr", ctypes.c_int32), ("inner_iters", ctypes.c_int32),
                ("steps_per_poll", ctypes.c_int32), ("use_graph", ctypes.c_int32),
                ("max_steps", ctypes.c_int32), ("verbose", ctypes.c_int32), ("profile", ctypes.c_int32),
                ("kernels", ctypes.c_int32), ("gr_rounds", ctypes.c_int32),
                ("gr_block_w", ctypes.c_int32), ("gr_block_h", ctypes.c_int32), ("gr_cap", ctypes.c_int32),
                ("pr_waves_per_cu", ctypes.c_int32), ("gr_waves_per_cu", ctypes.c_int32),
                ("lazy_caps", ctypes.c_int32), ("gr_cap_late", ctypes.c_int32)]


class Stats(ctypes.Structure):
    _fields_ = [("flow", ctypes.c_int64), ("pr_tiles", ctypes.c_int64), ("gr_tiles", ctypes.c_int64),
                ("pr_iters", ctypes.c_int64), ("steps", ctypes.c_int32), ("gr_rounds", ctypes.c_int32),
                ("ms", ctypes.c_double), ("pr_kernel_ms", ctypes.c_double), ("gr_kernel_ms", ctypes.c_double),
                ("sched_kernel_ms", ctypes.c_double), ("launches", ctypes.c_int64), ("gr_iters", ctypes.c_int64),
                ("gr_init_tiles", ctypes.c_int64), ("gr_exact_tiles", ctypes.c_int64), ("fin_items", ctypes.c_int64),
                ("build_kernel_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Instance(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("d_pos", _VP), ("d_neg", _VP),
                ("d_up", _VP), ("d_down", _VP), ("d_left", _VP), ("d_right", _VP), ("penalty", ctypes.c_int32)]


class Result(ctypes.Structure):
    _fields_ = [("request_id", ctypes.c_int64), ("session", ctypes.c_int32), ("status", ctypes.c_int32),
                ("flow", ctypes.c_int64), ("latency_ms", ctypes.c_double), ("solve_ms", ctypes.c_double),
                ("done_ms", ctypes.c_double), ("labels", ctypes.c_void_p), ("label_bytes", ctypes.c_size_t)]


_lib = None
_hip = None


def lib():
    """Load libgcfg.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = ctypes.CDLL(LIB_PATH)
        L.gcfg_last_error.restype = ctypes.c_char_p
        L.gcfg_now_ms.restype = ctypes.c_double
        L.gcfg_gen_mix_seed.restype = ctypes.c_uint64
        L.gcfg_gen_mix_seed.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        for name in ("gcfg_gen_ellipse_image", "gcfg_gen_perlin_image"):
            getattr(L, name).argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, _VP]
        L.gcfg_gen_seed_map.argtypes = [ctypes.c_int, ctypes.c_int, _VP]
        L.gcfg_gen_stress_tlinks.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, _VP, _VP]
        L.gcfg_gen_batch_shapes.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_double, ctypes.c_double, _VP,
                                            _VP]
        L.gcfg_dt_unquantized.argtypes = [_VP, _VP, _VP, _VP, _VP]
        L.gcfg_dt_get_beta.argtypes = [_VP, _I64P, _I64P, ctypes.POINTER(ctypes.c_double)]
        L.gcfg_server_create.argtypes = [ctypes.c_size_t, _VP, _VP]
        L.gcfg_server_submit_image.argtypes = [_VP, ctypes.c_int32, ctypes.c_int32, _VP, _VP, _VP, _I64P,
                                               ctypes.POINTER(ctypes.c_int32)]
        L.gcfg_server_submit_edit.argtypes = [_VP, ctypes.c_int32, ctypes.c_int32, _VP, ctypes.c_int32, _VP, _I64P]
        L.gcfg_server_load.argtypes = [_VP, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                       ctypes.POINTER(ctypes.c_size_t)]
        _lib = L
    return _lib


def _check(rc: int, what: str):
    if rc != GCFG_OK:
        raise GcfgError(rc, what, lib().gcfg_last_error().decode(errors="replace"))


def now_ms() -> float:
    return float(lib().gcfg_now_ms())


def mrf_load(path: str):
    """Middlebury 2-label MRF file -> (W, H, pos, neg, up, down, left, right) int32 [H, W] arrays."""
    W, H, L = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    _check(lib().gcfg_mrf_info(path.encode(), ctypes.byref(W), ctypes.byref(H), ctypes.byref(L)), "gcfg_mrf_info")
    arrs = [np.empty((H.value, W.value), np.int32) for _ in range(6)]
    _check(lib().gcfg_mrf_load(path.encode(), W.value, H.value, *[_VP(a.ctypes.data) for a in arrs]), "gcfg_mrf_load")
    return (W.value, H.value, *arrs)


def mrf_write(path: str, pos, neg, hcue, vcue, mult: int = 1):
    arrs = [np.ascontiguousarray(a, np.int32) for a in (pos, neg, hcue, vcue)]
    H, W = arrs[0].shape
    _check(lib().gcfg_mrf_write(path.encode(), W, H, int(mult), *[_VP(a.ctypes.data) for a in arrs]),
           "gcfg_mrf_write")


def dat_load(path: str):
    """Data-term file (Application.cu:15-24) -> (W, H, pos, neg) int32 [H, W]."""
    W, H = ctypes.c_int32(), ctypes.c_int32()
    _check(lib().gcfg_dat_info(path.encode(), ctypes.byref(W), ctypes.byref(H)), "gcfg_dat_info")
    pos = np.empty((H.value, W.value), np.int32)
    neg = np.empty((H.value, W.value), np.int32)
    _check(lib().gcfg_dat_load(path.encode(), W.value, H.value, _VP(pos.ctypes.data), _VP(neg.ctypes.data)),
           "gcfg_dat_load")
    return W.value, H.value, pos, neg


def dat_write(path: str, pos, neg):
    pos = np.ascontiguousarray(pos, np.int32)
    neg = np.ascontiguousarray(neg, np.int32)
    H, W = pos.shape
    _check(lib().gcfg_dat_write(path.encode(), W, H, _VP(pos.ctypes.data), _VP(neg.ctypes.data)), "gcfg_dat_write")


def mrf_write_tiled(path: str, out_w: int, out_h: int, pos, neg, penalty: int = 600):
    """writeRihannaMRFData (multiplyMRF.cpp:94-122) at out_w x out_h."""
    pos = np.ascontiguousarray(pos, np.int32)
    neg = np.ascontiguousarray(neg, np.int32)
    H, W = pos.shape
    _check(lib().gcfg_mrf_write_tiled(path.encode(), int(out_w), int(out_h), W, H, _VP(pos.ctypes.data),
                                      _VP(neg.ctypes.data), int(penalty)), "gcfg_mrf_write_tiled")


def read_ppm(path: str) -> np.ndarray:
    W, H = ctypes.c_int32(), ctypes.c_int32()
    _check(lib().gcfg_ppm_info(path.encode(), ctypes.byref(W), ctypes.byref(H)), "gcfg_ppm_info")
    img = np.empty((H.value, W.value, 3), np.uint8)
    _check(lib().gcfg_ppm_read(path.encode(), W.value, H.value, _VP(img.ctypes.data)), "gcfg_ppm_read")
    return img


def write_label_ppm(path: str, label, rgb=None):
    lab = np.ascontiguousarray(label, np.int32)
    H, W = lab.shape
    img = None if rgb is None else np.ascontiguousarray(rgb, np.uint8)
    _check(lib().gcfg_write_label_ppm(path.encode(), W, H, _VP(lab.ctypes.data),
                                      _VP(img.ctypes.data) if img is not None else None), "gcfg_write_label_ppm")


def write_label_pgm(path: str, label):
    lab = np.ascontiguousarray(label, np.int32)
    H, W = lab.shape
    _check(lib().gcfg_write_label_pgm(path.encode(), W, H, _VP(lab.ctypes.data)), "gcfg_write_label_pgm")


def tune_store(path: str, key: str, options: Options):
    _check(lib().gcfg_tune_store(path.encode(), key.encode(), ctypes.byref(options)), "gcfg_tune_store")


def tune_lookup(path: str, key: str):
    """Stored options for key, or None."""
    o = Options()
    rc = lib().gcfg_tune_lookup(path.encode(), key.encode(), ctypes.byref(o))
    return o if rc == GCFG_OK else None


def unpack_bits(words: np.ndarray, width: int, height: int) -> np.ndarray:
    """Packed label map (rows of ceil(W/64) uint64 words, bit x of word k =
    pixel 64 k + x) -> int32 labels [H, W]."""
    w = np.ascontiguousarray(words, dtype=np.uint64).reshape(height, -1)
    bits = np.unpackbits(w.view(np.uint8), axis=1, bitorder="little")
    return bits[:, :width].astype(np.int32)


def pack_bits(label: np.ndarray) -> np.ndarray:
    """int32 labels [H, W] -> packed words [H, ceil(W/64)] (inverse of unpack_bits)."""
    H, W = label.shape
    k = (W + 63) // 64
    padded = np.zeros((H, 64 * k), np.uint8)
    padded[:, :W] = label != 0
    return np.packbits(padded, axis=1, bitorder="little").view(np.uint64).reshape(H, k)


def options_dict(o: Options) -> dict:
    return {k: getattr(o, k) for k, _ in Options._fields_}


def default_options() -> Options:
    o = Options()
    lib().gcfg_default_options(ctypes.byref(o))
    return o


# ---------------------------------------------------------------------------
# HIP device memory (plumbing only).
def hip():
    global _hip
    if _hip is None:
        h = ctypes.CDLL("libamdhip64.so")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [_VP]
        h.hipMemcpy.argtypes = [_VP, _VP, ctypes.c_size_t, ctypes.c_int]
        h.hipMemset.argtypes = [_VP, ctypes.c_int, ctypes.c_size_t]
        h.hipGetErrorString.restype = ctypes.c_char_p
        h.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        h.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        h.hipStreamDestroy.argtypes = [_VP]
        h.hipStreamSynchronize.argtypes = [_VP]
        h.hipMemcpyAsync.argtypes = [_VP, _VP, ctypes.c_size_t, ctypes.c_int, _VP]
        _hip = h
    return _hip


def _hip_check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(rc).decode()}")


def gpu_available() -> bool:
    try:
        n = ctypes.c_int(0)
        return hip().hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


_pinned_keep = {}


def pinned_empty(shape, dtype) -> np.ndarray:
    """A numpy array in pinned host memory (hipHostMalloc), for inputs that
    the library copies to the device asynchronously. Freed with
    pinned_free(array) or at exit."""
    dt = np.dtype(dtype)
    n = int(np.prod(shape)) * dt.itemsize
    p = ctypes.c_void_p()
    h = hip()
    h.hipHostMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t, ctypes.c_uint]
    _hip_check(h.hipHostMalloc(ctypes.byref(p), max(n, 1), 0), "hipHostMalloc")
    buf = (ctypes.c_uint8 * max(n, 1)).from_address(p.value)
    a = np.frombuffer(buf, dtype=np.uint8, count=n).view(dt).reshape(shape)
    _pinned_keep[a.ctypes.data] = p.value
    return a


def pinned_free(a: np.ndarray):
    p = _pinned_keep.pop(a.ctypes.data, None)
    if p:
        hip().hipHostFree.argtypes = [_VP]
        hip().hipHostFree(_VP(p))


def device_synchronize():
    _hip_check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Stream:
    """A HIP stream (``ptr`` is the hipStream_t) for ordering against a caller's work."""

    def __init__(self):
        p = ctypes.c_void_p()
        _hip_check(hip().hipStreamCreate(ctypes.byref(p)), "hipStreamCreate")
        self.ptr = p.value

    def synchronize(self):
        _hip_check(hip().hipStreamSynchronize(_VP(self.ptr)), "hipStreamSynchronize")

    def copy_async(self, dst_ptr: int, src_ptr: int, nbytes: int, kind: int = 3):
        """hipMemcpyAsync on this stream (kind 3: device to device)."""
        _hip_check(hip().hipMemcpyAsync(_VP(dst_ptr), _VP(src_ptr), int(nbytes), int(kind), _VP(self.ptr)),
                   "hipMemcpyAsync")

    def close(self):
        if self.ptr:
            hip().hipStreamDestroy(_VP(self.ptr))
            self.ptr = None


class DeviceBuffer:
    """A hipMalloc'd buffer; ``ptr`` is the device address."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _hip_check(hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 16)), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        _hip_check(hip().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")

    def download(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        assert out.nbytes <= self.nbytes
        _hip_check(hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2), "hipMemcpy D2H")
        return out

    def copy_from_device(self, src_ptr: int, nbytes: int):
        assert nbytes <= self.nbytes
        _hip_check(hip().hipMemcpy(self.ptr, src_ptr, nbytes, 3), "hipMemcpy D2D")

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Seeded generators (host-only).
def gen_ellipse_image(width: int, height: int, seed: int) -> np.ndarray:
    img = np.empty((height, width, 3), np.uint8)
    _check(lib().gcfg_gen_ellipse_image(width, height, seed, img.ctypes.data), "gcfg_gen_ellipse_image")
    return img


def gen_perlin_image(width: int, height: int, seed: int) -> np.ndarray:
    img = np.empty((height, width, 3), np.uint8)
    _check(lib().gcfg_gen_perlin_image(width, height, seed, img.ctypes.data), "gcfg_gen_perlin_image")
    return img


def gen_seed_map(width: int, height: int) -> np.ndarray:
    s = np.empty((height, width), np.uint8)
    _check(lib().gcfg_gen_seed_map(width, height, s.ctypes.data), "gcfg_gen_seed_map")
    return s


def gen_stress_tlinks(width: int, height: int, seed: int):
    pos = np.empty((height, width), np.int32)
    neg = np.empty((height, width), np.int32)
    _check(lib().gcfg_gen_stress_tlinks(width, height, seed, pos.ctypes.data, neg.ctypes.data),
           "gcfg_gen_stress_tlinks")
    return pos, neg


def gen_batch_shapes(count: int, seed: int, area_min: float, area_max: float):
    w = np.empty(count, np.int32)
    h = np.empty(count, np.int32)
    _check(lib().gcfg_gen_batch_shapes(count, seed, area_min, area_max, w.ctypes.data, h.ctypes.data),
           "gcfg_gen_batch_shapes")
    return w, h


def mix_seed(seed: int, stream: int) -> int:
    return int(lib().gcfg_gen_mix_seed(seed, stream))


# ---------------------------------------------------------------------------
class Graph:
    """The GC_* drop-in: one W x H grid graph on the device."""

    def __init__(self, width: int, height: int, options: Options | None = None, neighborhood: int = 4):
        self.width, self.height = int(width), int(height)
        if neighborhood not in (4, 8):
            raise ValueError("neighborhood must be 4 or 8")
        self.neighborhood = neighborhood
        h = ctypes.c_void_p()
        create = lib().gcfg_graph_create_n8 if neighborhood == 8 else lib().gcfg_graph_create
        _check(create(self.width, self.height, ctypes.byref(h)), "gcfg_graph_create")
        self.h = h.value
        self._keep = {}
        if options is not None:
            self.set_options(options)

    def set_options(self, options: Options):
        _check(lib().gcfg_graph_set_options(_VP(self.h), ctypes.byref(options)), "gcfg_graph_set_options")

    def _dev(self, name, a):
        if isinstance(a, DeviceBuffer):
            self._keep[name] = a
            return a.ptr
        if isinstance(a, int):
            return a
        buf = DeviceBuffer.from_array(np.ascontiguousarray(a, dtype=np.int32))
        self._keep[name] = buf
        return buf.ptr

    def set_dataterms(self, pos, neg):
        _check(lib().gcfg_graph_set_dataterms(_VP(self.h), _VP(self._dev("pos", pos)), _VP(self._dev("neg", neg))),
               "gcfg_graph_set_dataterms")

    def set_edges(self, up, down, left, right):
        ptrs = [_VP(self._dev(n, a)) for n, a in (("up", up), ("down", down), ("left", left), ("right", right))]
        _check(lib().gcfg_graph_set_edges(_VP(self.h), *ptrs), "gcfg_graph_set_edges")

    def set_edges8(self, up, down, left, right, upleft, upright, downleft, downright):
        """GC_SetEdges of an 8-neighbourhood graph (GraphCut.cu:271-289)."""
        names = ("up", "down", "left", "right", "upleft", "upright", "downleft", "downright")
        arrs = (up, down, left, right, upleft, upright, downleft, downright)
        ptrs = [_VP(self._dev(n, a)) for n, a in zip(names, arrs)]
        _check(lib().gcfg_graph_set_edges8(_VP(self.h), *ptrs), "gcfg_graph_set_edges8")

    def set_penalty(self, penalty: int):
        _check(lib().gcfg_graph_set_penalty(_VP(self.h), int(penalty)), "gcfg_graph_set_penalty")

    def build(self):
        _check(lib().gcfg_graph_build(_VP(self.h)), "gcfg_graph_build")

    def optimize(self):
        label = np.empty((self.height, self.width), np.int32)
        st = Stats()
        _check(lib().gcfg_graph_optimize(_VP(self.h), _VP(label.ctypes.data), ctypes.byref(st)), "gcfg_graph_optimize")
        return label, st

    def optimize_device(self, d_label: DeviceBuffer):
        st = Stats()
        _check(lib().gcfg_graph_optimize_device(_VP(self.h), _VP(d_label.ptr), ctypes.byref(st)),
               "gcfg_graph_optimize_device")
        return st

    def autotune(self, trials: int = 2):
        """Pick solver options by timing cold solves of this graph (built inputs)."""
        best, ms = Options(), ctypes.c_double()
        _check(lib().gcfg_graph_autotune(_VP(self.h), int(trials), ctypes.byref(best), ctypes.byref(ms)),
               "gcfg_graph_autotune")
        return best, float(ms.value)

    def read_distances(self):
        dist = np.empty((self.height, self.width), np.int32)
        _check(lib().gcfg_graph_read_distances(_VP(self.h), _VP(dist.ctypes.data)), "gcfg_graph_read_distances")
        return dist

    def read_state(self):
        caps = np.empty((self.height, self.width, 4), np.int32)
        exc = np.empty((self.height, self.width), np.int32)
        _check(lib().gcfg_graph_read_state(_VP(self.h), _VP(caps.ctypes.data), _VP(exc.ctypes.data)),
               "gcfg_graph_read_state")
        return caps, exc

    def read_diag_state(self):
        """Diagonal residual capacities (up-left, up-right, down-left, down-right) of an 8-neighbourhood graph."""
        caps8 = np.empty((self.height, self.width, 4), np.int32)
        _check(lib().gcfg_graph_read_diag_state(_VP(self.h), _VP(caps8.ctypes.data)), "gcfg_graph_read_diag_state")
        return caps8

    def set_stream(self, stream_ptr: int):
        """Order this handle's work against a caller HIP stream (0: none)."""
        _check(lib().gcfg_graph_set_stream(_VP(self.h), _VP(stream_ptr or None)), "gcfg_graph_set_stream")

    def close(self):
        if self.h:
            lib().gcfg_graph_destroy(_VP(self.h))
            self.h = None
        self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve(width, height, pos, neg, edges=None, penalty=0, options: Options | None = None):
    """One plain drop-in solve. Returns (flow, labels[H, W], stats)."""
    g = Graph(width, height, options)
    try:
        g.set_dataterms(pos, neg)
        if edges is not None:
            g.set_edges(*edges)
        else:
            g.set_penalty(penalty)
        g.build()
        label, st = g.optimize()
        return int(st.flow), label, st
    finally:
        g.close()


def solve8(width, height, pos, neg, dirs8=None, penalty=0, options: Options | None = None):
    """One plain 8-neighbourhood solve. ``dirs8`` = (left, right, up, down,
    upleft, upright, downleft, downright) arrays or None for a uniform
    ``penalty``. Returns (flow, labels[H, W], stats)."""
    g = Graph(width, height, options, neighborhood=8)
    try:
        g.set_dataterms(pos, neg)
        if dirs8 is not None:
            left, right, up, down, ul, ur, dl, dr = dirs8
            g.set_edges8(up, down, left, right, ul, ur, dl, dr)
        else:
            g.set_penalty(penalty)
        g.build()
        label, st = g.optimize()
        return int(st.flow), label, st
    finally:
        g.close()


# ---------------------------------------------------------------------------
class DataTerm:
    """The DataTerm.h stage: seed histograms, t-links and n-links on the device."""

    def __init__(self, width: int, height: int):
        self.width, self.height = int(width), int(height)
        h = ctypes.c_void_p()
        _check(lib().gcfg_dt_create(self.width, self.height, ctypes.byref(h)), "gcfg_dt_create")
        self.h = h.value

    def set_image(self, rgb: np.ndarray):
        rgb = np.ascontiguousarray(rgb, dtype=np.uint8)
        assert rgb.size == 3 * self.width * self.height
        self._img = rgb
        _check(lib().gcfg_dt_set_image(_VP(self.h), _VP(rgb.ctypes.data)), "gcfg_dt_set_image")

    def set_seeds(self, seeds: np.ndarray):
        seeds = np.ascontiguousarray(seeds, dtype=np.uint8)
        self._seeds = seeds
        _check(lib().gcfg_dt_set_seeds(_VP(self.h), _VP(seeds.ctypes.data)), "gcfg_dt_set_seeds")

    def add_seed_rect(self, x0, y0, w, h, cls):
        _check(lib().gcfg_dt_add_seed_rect(_VP(self.h), x0, y0, w, h, cls), "gcfg_dt_add_seed_rect")

    def reset(self):
        _check(lib().gcfg_dt_reset(_VP(self.h)), "gcfg_dt_reset")

    def set_lambda(self, lam: int):
        _check(lib().gcfg_dt_set_lambda(_VP(self.h), int(lam)), "gcfg_dt_set_lambda")

    def compute(self):
        _check(lib().gcfg_dt_compute(_VP(self.h)), "gcfg_dt_compute")

    def set_model(self, model: int):
        """DT_MODEL_HISTOGRAM (default) or DT_MODEL_KDE (the reference's kernel density)."""
        _check(lib().gcfg_dt_set_model(_VP(self.h), int(model)), "gcfg_dt_set_model")

    def set_sigma(self, sigma: int):
        """setSigmaGaussian (DataTerm.cpp:97-103)."""
        _check(lib().gcfg_dt_set_sigma(_VP(self.h), int(sigma)), "gcfg_dt_set_sigma")

    def kde_select(self, x: int, y: int):
        """selectOnMap (DataTerm.cpp:105-127): the 5 x 5 square's colours join the density."""
        _check(lib().gcfg_dt_kde_select(_VP(self.h), int(x), int(y)), "gcfg_dt_kde_select")

    def kde_count(self) -> int:
        n = ctypes.c_int32(0)
        _check(lib().gcfg_dt_kde_count(_VP(self.h), ctypes.byref(n)), "gcfg_dt_kde_count")
        return int(n.value)

    def link_ptrs(self):
        ps = [ctypes.c_void_p() for _ in range(6)]
        _check(lib().gcfg_dt_get_links(_VP(self.h), *[ctypes.byref(p) for p in ps]), "gcfg_dt_get_links")
        return [p.value for p in ps]

    def links(self):
        """Download pos, neg, up, down, left, right as int32 [H, W] arrays."""
        n = self.width * self.height
        out = []
        for p in self.link_ptrs():
            a = np.empty(n, np.int32)
            _hip_check(hip().hipMemcpy(a.ctypes.data, p, a.nbytes, 2), "hipMemcpy D2H")
            out.append(a.reshape(self.height, self.width))
        return out

    def histograms(self):
        h = np.empty(96, np.int32)
        c = np.empty(2, np.int32)
        _check(lib().gcfg_dt_get_histograms(_VP(self.h), _VP(h.ctypes.data), _VP(c.ctypes.data)),
               "gcfg_dt_get_histograms")
        return h.reshape(2, 3, 16), c

    def beta(self):
        s, p, b = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_double()
        _check(lib().gcfg_dt_get_beta(_VP(self.h), ctypes.byref(s), ctypes.byref(p), ctypes.byref(b)),
               "gcfg_dt_get_beta")
        return int(s.value), int(p.value), float(b.value)

    def unquantized(self):
        n = self.width * self.height
        outs = [np.empty(n, np.float64) for _ in range(4)]
        _check(lib().gcfg_dt_unquantized(_VP(self.h), *[_VP(o.ctypes.data) for o in outs]), "gcfg_dt_unquantized")
        return [o.reshape(self.height, self.width) for o in outs]

    def set_stream(self, stream_ptr: int):
        """Order this handle's work against a caller HIP stream (0: none)."""
        _check(lib().gcfg_dt_set_stream(_VP(self.h), _VP(stream_ptr or None)), "gcfg_dt_set_stream")

    def close(self):
        if self.h:
            lib().gcfg_dt_destroy(_VP(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DtImage(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("d_rgb", _VP), ("d_seeds", _VP),
                ("d_pos", _VP), ("d_neg", _VP), ("d_up", _VP), ("d_down", _VP), ("d_left", _VP), ("d_right", _VP)]


class DataTermBatch:
    """Histogram data term of many images in one launch per pass
    (gcfg_dt_batch_*). ``images`` = list of (rgb [H, W, 3] u8, seeds [H, W] u8
    or None); the device planes of image i are ``planes[i]`` (pos, neg, up,
    down, left, right DeviceBuffers)."""

    def __init__(self, images, lam: int = 50):
        n = len(images)
        arr = (DtImage * n)()
        self.inputs, self.planes, self.shapes = [], [], []
        for i, (rgb, seeds) in enumerate(images):
            rgb = np.ascontiguousarray(rgb, np.uint8)
            H, W = rgb.shape[:2]
            drgb = DeviceBuffer.from_array(rgb)
            dseeds = DeviceBuffer.from_array(np.ascontiguousarray(seeds, np.uint8)) if seeds is not None else None
            outs = [DeviceBuffer(4 * W * H) for _ in range(6)]
            self.inputs.append((drgb, dseeds))
            self.planes.append(outs)
            self.shapes.append((W, H))
            arr[i].width, arr[i].height = W, H
            arr[i].d_rgb = drgb.ptr
            arr[i].d_seeds = dseeds.ptr if dseeds is not None else None
            arr[i].d_pos, arr[i].d_neg, arr[i].d_up, arr[i].d_down, arr[i].d_left, arr[i].d_right = \
                (o.ptr for o in outs)
        h = ctypes.c_void_p()
        _check(lib().gcfg_dt_batch_create(n, arr, int(lam), ctypes.byref(h)), "gcfg_dt_batch_create")
        self.h = h.value

    def compute(self):
        _check(lib().gcfg_dt_batch_compute(_VP(self.h)), "gcfg_dt_batch_compute")

    def links(self, i: int):
        W, H = self.shapes[i]
        return [b.download(np.int32, W * H).reshape(H, W) for b in self.planes[i]]

    def close(self):
        if self.h:
            lib().gcfg_dt_batch_destroy(_VP(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
@dataclass
class BatchItem:
    width: int
    height: int
    pos: np.ndarray
    neg: np.ndarray
    edges: tuple | None = None
    penalty: int = 0


class Batch:
    """Many instances of uneven size solved together on one engine."""

    def __init__(self, items: list[BatchItem], options: Options | None = None):
        self.items = items
        self._bufs = []
        arr = (Instance * len(items))()
        def dev(v):
            # DeviceBuffer (borrowed, kept alive here) or host array (uploaded)
            if isinstance(v, DeviceBuffer):
                return v
            return DeviceBuffer.from_array(np.ascontiguousarray(v, dtype=np.int32))

        for i, it in enumerate(items):
            d = {"pos": dev(it.pos), "neg": dev(it.neg)}
            if it.edges is not None:
                for k, v in zip(("up", "down", "left", "right"), it.edges):
                    d[k] = dev(v)
            self._bufs.append(d)
            arr[i].width, arr[i].height, arr[i].penalty = it.width, it.height, it.penalty
            arr[i].d_pos, arr[i].d_neg = d["pos"].ptr, d["neg"].ptr
            if it.edges is not None:
                arr[i].d_up, arr[i].d_down, arr[i].d_left, arr[i].d_right = (d[k].ptr for k in
                                                                               ("up", "down", "left", "right"))
        h = ctypes.c_void_p()
        _check(lib().gcfg_batch_create(len(items), arr, ctypes.byref(options) if options is not None else None,
                                       ctypes.byref(h)), "gcfg_batch_create")
        self.h = h.value
        self.labels = [DeviceBuffer(4 * it.width * it.height) for it in items]

    def solve(self, labels_to_host: bool = True):
        n = len(self.items)
        flows = np.zeros(n, np.int64)
        st = Stats()
        if labels_to_host:
            host = [np.empty((it.height, it.width), np.int32) for it in self.items]
            ptrs = (ctypes.c_void_p * n)(*[a.ctypes.data for a in host])
            _check(lib().gcfg_batch_solve(_VP(self.h), ptrs, _VP(flows.ctypes.data), ctypes.byref(st)),
                   "gcfg_batch_solve")
            return flows, host, st
        ptrs = (ctypes.c_void_p * n)(*[b.ptr for b in self.labels])
        _check(lib().gcfg_batch_solve_device(_VP(self.h), ptrs, _VP(flows.ctypes.data), ctypes.byref(st)),
               "gcfg_batch_solve_device")
        return flows, None, st

    def solve_bits(self, copy: bool = True):
        """Cold solve with the labels on the host packed one bit per pixel
        (gcfg_batch_solve_bits). Returns (flows, [uint64 array [H, ceil(W/64)]
        per instance], stats); with copy=False the arrays view the batch's
        pinned buffer, valid until the next solve."""
        n = len(self.items)
        flows = np.zeros(n, np.int64)
        st = Stats()
        ptrs = (ctypes.POINTER(ctypes.c_uint64) * n)()
        _check(lib().gcfg_batch_solve_bits(_VP(self.h), ptrs, _VP(flows.ctypes.data), ctypes.byref(st)),
               "gcfg_batch_solve_bits")
        out = []
        for it, p in zip(self.items, ptrs):
            a = np.ctypeslib.as_array(p, shape=(it.height, (it.width + 63) // 64))
            out.append(a.copy() if copy else a)
        return flows, out, st

    def autotune(self, trials: int = 2):
        """Pick solver options by timing cold solves of the whole batch."""
        best, ms = Options(), ctypes.c_double()
        _check(lib().gcfg_batch_autotune(_VP(self.h), int(trials), ctypes.byref(best), ctypes.byref(ms)),
               "gcfg_batch_autotune")
        return best, float(ms.value)

    def device_labels(self, i: int) -> np.ndarray:
        it = self.items[i]
        return self.labels[i].download(np.int32, it.width * it.height).reshape(it.height, it.width)

    def set_stream(self, stream_ptr: int):
        """Order this handle's work against a caller HIP stream (0: none)."""
        _check(lib().gcfg_batch_set_stream(_VP(self.h), _VP(stream_ptr or None)), "gcfg_batch_set_stream")

    def close(self):
        if self.h:
            lib().gcfg_batch_destroy(_VP(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
class Server:
    """Continuous-batching segmentation server with device-resident sessions."""

    def __init__(self, hbm_budget_bytes: int, options: Options | None = None, label_format: int = LABELS_INT32):
        h = ctypes.c_void_p()
        _check(lib().gcfg_server_create(int(hbm_budget_bytes), ctypes.byref(options) if options is not None else None,
                                        ctypes.byref(h)), "gcfg_server_create")
        self.h = h.value
        self._pending = {}
        self._lent = {}  # request id -> (shape, dtype) of labels the server lends (borrow=True)
        self.label_format = LABELS_INT32
        if label_format != LABELS_INT32:
            self.set_label_format(label_format)

    def set_label_format(self, fmt: int):
        """LABELS_INT32 (int32 [H, W]) or LABELS_BITS (uint64 [H, ceil(W/64)]) for later requests."""
        _check(lib().gcfg_server_set_label_format(_VP(self.h), int(fmt)), "gcfg_server_set_label_format")
        self.label_format = fmt

    def _label_shape(self, H, W):
        if self.label_format == LABELS_BITS:
            return (H, (W + 63) // 64), np.uint64
        return (H, W), np.int32

    def _label(self, H, W, borrow):
        """The caller's label array, or None when the result lends the labels."""
        if borrow:
            return None
        shape, dt = self._label_shape(H, W)
        return np.empty(shape, dt)

    def _lend(self, rid, H, W, borrow):
        if borrow:
            self._lent[rid] = self._label_shape(H, W)

    def submit_image(self, rgb: np.ndarray, seeds: np.ndarray, borrow: bool = False):
        """borrow=True: no label array; the result's "labels" is a view of the
        server's pinned memory, valid until the next poll."""
        rgb = np.ascontiguousarray(rgb, dtype=np.uint8)
        seeds = np.ascontiguousarray(seeds, dtype=np.uint8)
        H, W = seeds.shape
        label = self._label(H, W, borrow)
        rid, sess = ctypes.c_int64(), ctypes.c_int32()
        _check(lib().gcfg_server_submit_image(_VP(self.h), W, H, _VP(rgb.ctypes.data), _VP(seeds.ctypes.data),
                                              _VP(label.ctypes.data) if label is not None else None,
                                              ctypes.byref(rid), ctypes.byref(sess)), "gcfg_server_submit_image")
        self._pending[rid.value] = (rgb, seeds, label)
        self._lend(rid.value, H, W, borrow)
        return rid.value, sess.value, label

    def submit_edit(self, session: int, strokes, delta: int = 1 << 16, shape=None, borrow: bool = False):
        st = np.ascontiguousarray(np.asarray(strokes, dtype=np.int32).reshape(-1, 5))
        if shape is None:
            raise ValueError("shape (H, W) of the session image is required for the label buffer")
        label = self._label(*shape, borrow)
        rid = ctypes.c_int64()
        _check(lib().gcfg_server_submit_edit(_VP(self.h), int(session), st.shape[0], _VP(st.ctypes.data), int(delta),
                                             _VP(label.ctypes.data) if label is not None else None,
                                             ctypes.byref(rid)), "gcfg_server_submit_edit")
        self._pending[rid.value] = (st, label)
        self._lend(rid.value, *shape, borrow)
        return rid.value, label

    def submit_frame(self, session: int, rgb: np.ndarray):
        """Next video frame of a session (same W x H): warm start from its residual graph."""
        rgb = np.ascontiguousarray(rgb, dtype=np.uint8)
        H, W = rgb.shape[:2]
        label = self._label(H, W, False)
        rid = ctypes.c_int64()
        _check(lib().gcfg_server_submit_frame(_VP(self.h), int(session), W, H, _VP(rgb.ctypes.data),
                                              _VP(label.ctypes.data), ctypes.byref(rid)), "gcfg_server_submit_frame")
        self._pending[rid.value] = (rgb, label)
        return rid.value, label

    def poll(self, wait: bool = False, max_results: int = 256):
        res = (Result * max_results)()
        n = ctypes.c_int32()
        _check(lib().gcfg_server_poll(_VP(self.h), int(wait), max_results, res, ctypes.byref(n)), "gcfg_server_poll")
        out = []
        for i in range(n.value):
            r = res[i]
            self._pending.pop(r.request_id, None)
            d = {k: getattr(r, k) for k, _ in Result._fields_ if k not in ("labels", "label_bytes")}
            lent = self._lent.pop(r.request_id, None)
            if lent is not None and r.labels:
                shape, dt = lent
                buf = (ctypes.c_uint8 * r.label_bytes).from_address(r.labels)
                d["labels"] = np.frombuffer(buf, dtype=dt, count=int(np.prod(shape))).reshape(shape)
            out.append(d)
        return out

    def load(self):
        a, q, b = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_size_t()
        _check(lib().gcfg_server_load(_VP(self.h), ctypes.byref(a), ctypes.byref(q), ctypes.byref(b)),
               "gcfg_server_load")
        return a.value, q.value, b.value

    def close_session(self, session: int):
        _check(lib().gcfg_server_close_session(_VP(self.h), int(session)), "gcfg_server_close_session")

    def close(self):
        if self.h:
            lib().gcfg_server_destroy(_VP(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
