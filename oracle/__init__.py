"""CPU oracle for the binary-MRF min-cut path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg import this module, and only as the checker (or, for the
baseline, as the timed CPU restatement of the reference schedule). The product
library in ``gc-fg_amd/`` never imports, links or calls it.

Functions (numpy in / numpy out, ctypes over ``oracle/liboracle.so``):

* :func:`maxflow` -- exact int64 max-flow and the "cannot reach the sink"
  label map on the graph of ``GraphCutKernels.cu:112-214`` (interior nodes
  only, arcs into the border zeroed); labels as ``GraphCutKernels.cu:228-351``
  with ``SPREAD_ZEROS`` (``GraphCut.h:37``).
* :func:`cut_value` -- cut value of a labelling on the same graph.
* :func:`maxflow8` / :func:`cut_value8` -- the same on the 8-neighbourhood
  grid (``NEIGHBORHOOD == 8``, ``GraphCut.h:22,65-70``).
* :func:`dataterm` -- per-pixel double-precision colour data term (16-bin
  seed histograms, ``round(64 * -ln p)`` t-links, ``round(lambda *
  exp(-d^2 / 2 beta))`` n-links), the stage that sits behind ``DataTerm.h``.
* :func:`dataterm_kde` -- the reference's own Gaussian kernel-density colour
  model (``DataTermKernels.cu:19-150``) as a per-pixel double sum over the
  selected seed colours.
* :func:`refpr` -- CPU restatement of the reference push-relabel schedule
  (``GraphCut.cu:385-534``), used as the timed CPU baseline.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_I32P = ctypes.POINTER(ctypes.c_int32)
_I64P = ctypes.POINTER(ctypes.c_int64)
_U8P = ctypes.POINTER(ctypes.c_uint8)
_F64P = ctypes.POINTER(ctypes.c_double)


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B", "-C", _HERE], check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        lib = ctypes.CDLL(_LIB_PATH)
        lib.oracle_maxflow_grid.restype = ctypes.c_int
        lib.oracle_maxflow_grid.argtypes = [ctypes.c_int, ctypes.c_int, _I32P, _I32P, _I32P, _I32P, _I32P, _I32P,
                                            ctypes.c_int32, _I32P, _I64P]
        lib.oracle_cut_value.restype = ctypes.c_int64
        lib.oracle_cut_value.argtypes = [ctypes.c_int, ctypes.c_int, _I32P, _I32P, _I32P, _I32P, _I32P, _I32P,
                                         ctypes.c_int32, _I32P]
        lib.oracle_dataterm.restype = ctypes.c_int
        lib.oracle_dataterm.argtypes = [ctypes.c_int, ctypes.c_int, _U8P, _U8P, ctypes.c_int32, _I32P, _I32P,
                                        _I64P, _I64P, _F64P, _I32P, _I32P, _I32P, _I32P, _I32P, _I32P, _F64P,
                                        _F64P, _F64P, _F64P]
        lib.oracle_maxflow_grid8.restype = ctypes.c_int
        lib.oracle_maxflow_grid8.argtypes = [ctypes.c_int, ctypes.c_int, _I32P, _I32P, ctypes.c_void_p,
                                             ctypes.c_int32, _I32P, _I64P]
        lib.oracle_cut_value8.restype = ctypes.c_int64
        lib.oracle_cut_value8.argtypes = [ctypes.c_int, ctypes.c_int, _I32P, _I32P, ctypes.c_void_p,
                                          ctypes.c_int32, _I32P]
        lib.oracle_maxflow_grid_state.restype = ctypes.c_int
        lib.oracle_maxflow_grid_state.argtypes = [ctypes.c_int, ctypes.c_int, _I32P, _I32P, _I32P, _I32P, _I32P,
                                                  _I32P, ctypes.c_int32, _I32P, _I64P, _I32P, _I32P]
        lib.oracle_certify_grid.restype = ctypes.c_int
        lib.oracle_certify_grid.argtypes = [ctypes.c_int, ctypes.c_int, _I32P, _I32P, _I32P, _I32P, _I32P, _I32P,
                                            ctypes.c_int32, _I32P, _I32P, _I32P, ctypes.c_int64]
        lib.oracle_kde_gauss.restype = None
        lib.oracle_kde_gauss.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
        lib.oracle_dataterm_kde.restype = ctypes.c_int
        lib.oracle_dataterm_kde.argtypes = [ctypes.c_int, ctypes.c_int, _U8P, ctypes.c_int, _U8P, ctypes.c_int,
                                            _I32P, _I32P, _F64P, _F64P, _F64P]
        lib.oracle_dataterm_model.restype = ctypes.c_int
        lib.oracle_dataterm_model.argtypes = [ctypes.c_int, ctypes.c_int, _U8P, _U8P, _U8P, ctypes.c_int32, _I32P,
                                              _I32P, _I64P, _I64P, _F64P, _I32P, _I32P, _I32P, _I32P, _I32P, _I32P,
                                              _F64P, _F64P, _F64P, _F64P]
        lib.oracle_refpr_solve.restype = ctypes.c_int
        lib.oracle_refpr_solve.argtypes = [ctypes.c_int, ctypes.c_int, _I32P, _I32P, _I32P, _I32P, _I32P, _I32P,
                                           ctypes.c_int32, ctypes.c_int, _I32P, _I64P, _I64P]
        _lib = lib
    return _lib


def _ptr(a, ctype):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def _i32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


def _edges(edges):
    if edges is None:
        return (None,) * 4
    up, down, left, right = edges
    return tuple(_i32(e).ravel() for e in (up, down, left, right))


def maxflow(width, height, pos, neg, edges=None, penalty=0):
    """Exact max-flow. ``edges`` = (up, down, left, right) arrays or None for a
    uniform ``penalty``. Returns (flow:int, label:int32[H, W])."""
    lib = _load()
    pos, neg = _i32(pos).ravel(), _i32(neg).ravel()
    up, down, left, right = _edges(edges)
    label = np.zeros(width * height, np.int32)
    flow = ctypes.c_int64(0)
    rc = lib.oracle_maxflow_grid(width, height, _ptr(pos, ctypes.c_int32), _ptr(neg, ctypes.c_int32),
                                 _ptr(up, ctypes.c_int32), _ptr(down, ctypes.c_int32), _ptr(left, ctypes.c_int32),
                                 _ptr(right, ctypes.c_int32), int(penalty), _ptr(label, ctypes.c_int32),
                                 ctypes.byref(flow))
    if rc != 0:
        raise RuntimeError(f"oracle_maxflow_grid failed: {rc}")
    return int(flow.value), label.reshape(height, width)


def maxflow_state(width, height, pos, neg, edges=None, penalty=0):
    """maxflow plus the final residual state: (flow, label, caps[H, W, 4], excess[H, W])."""
    lib = _load()
    pos, neg = _i32(pos).ravel(), _i32(neg).ravel()
    up, down, left, right = _edges(edges)
    label = np.zeros(width * height, np.int32)
    caps = np.zeros(width * height * 4, np.int32)
    exc = np.zeros(width * height, np.int32)
    flow = ctypes.c_int64(0)
    rc = lib.oracle_maxflow_grid_state(width, height, _ptr(pos, ctypes.c_int32), _ptr(neg, ctypes.c_int32),
                                       _ptr(up, ctypes.c_int32), _ptr(down, ctypes.c_int32),
                                       _ptr(left, ctypes.c_int32), _ptr(right, ctypes.c_int32), int(penalty),
                                       _ptr(label, ctypes.c_int32), ctypes.byref(flow), _ptr(caps, ctypes.c_int32),
                                       _ptr(exc, ctypes.c_int32))
    if rc != 0:
        raise RuntimeError(f"oracle_maxflow_grid_state failed: {rc}")
    return (int(flow.value), label.reshape(height, width), caps.reshape(height, width, 4),
            exc.reshape(height, width))


def cut_value(width, height, pos, neg, label, edges=None, penalty=0) -> int:
    lib = _load()
    pos, neg = _i32(pos).ravel(), _i32(neg).ravel()
    up, down, left, right = _edges(edges)
    label = _i32(label).ravel()
    return int(lib.oracle_cut_value(width, height, _ptr(pos, ctypes.c_int32), _ptr(neg, ctypes.c_int32),
                                    _ptr(up, ctypes.c_int32), _ptr(down, ctypes.c_int32),
                                    _ptr(left, ctypes.c_int32), _ptr(right, ctypes.c_int32), int(penalty),
                                    _ptr(label, ctypes.c_int32)))


def certify(width, height, pos, neg, caps, excess, label, flow, edges=None, penalty=0) -> int:
    """Optimality certificate of a solver's residual state (caps H x W x 4 in
    left, right, up, down order; excess; labels; flow). 0 if it holds, else
    the number of the first failing check (1 feasibility, 2 conservation,
    3 labels, 4 flow / cut value)."""
    lib = _load()
    pos, neg = _i32(pos).ravel(), _i32(neg).ravel()
    up, down, left, right = _edges(edges)
    caps, exc, lab = _i32(caps).ravel(), _i32(excess).ravel(), _i32(label).ravel()
    return int(lib.oracle_certify_grid(width, height, _ptr(pos, ctypes.c_int32), _ptr(neg, ctypes.c_int32),
                                       _ptr(up, ctypes.c_int32), _ptr(down, ctypes.c_int32),
                                       _ptr(left, ctypes.c_int32), _ptr(right, ctypes.c_int32), int(penalty),
                                       _ptr(caps, ctypes.c_int32), _ptr(exc, ctypes.c_int32),
                                       _ptr(lab, ctypes.c_int32), int(flow)))


DIRS8 = ("left", "right", "up", "down", "upleft", "upright", "downleft", "downright")


def _planes8(dirs8):
    """(keep-alive arrays, void* to an array of 8 plane pointers) or (None, None)."""
    if dirs8 is None:
        return None, None
    arrs = [_i32(a).ravel() for a in dirs8]
    assert len(arrs) == 8
    ptrs = (ctypes.c_void_p * 8)(*[a.ctypes.data for a in arrs])
    return (arrs, ptrs), ctypes.cast(ptrs, ctypes.c_void_p)


def maxflow8(width, height, pos, neg, dirs8=None, penalty=0):
    """Exact max-flow on the 8-neighbourhood grid. ``dirs8`` = eight arrays in
    the order of DIRS8 or None for a uniform ``penalty``. Returns (flow, label)."""
    lib = _load()
    pos, neg = _i32(pos).ravel(), _i32(neg).ravel()
    keep, pp = _planes8(dirs8)
    label = np.zeros(width * height, np.int32)
    flow = ctypes.c_int64(0)
    rc = lib.oracle_maxflow_grid8(width, height, _ptr(pos, ctypes.c_int32), _ptr(neg, ctypes.c_int32), pp,
                                  int(penalty), _ptr(label, ctypes.c_int32), ctypes.byref(flow))
    del keep
    if rc != 0:
        raise RuntimeError(f"oracle_maxflow_grid8 failed: {rc}")
    return int(flow.value), label.reshape(height, width)


def cut_value8(width, height, pos, neg, label, dirs8=None, penalty=0) -> int:
    lib = _load()
    pos, neg = _i32(pos).ravel(), _i32(neg).ravel()
    keep, pp = _planes8(dirs8)
    label = _i32(label).ravel()
    v = int(lib.oracle_cut_value8(width, height, _ptr(pos, ctypes.c_int32), _ptr(neg, ctypes.c_int32), pp,
                                  int(penalty), _ptr(label, ctypes.c_int32)))
    del keep
    return v


def dataterm(width, height, rgb, seeds, lam=50, model_rgb=None):
    """Per-pixel double-precision data term. Returns a dict of numpy arrays.
    ``model_rgb``: take the seed histograms from this image instead (a video
    session's first frame)."""
    lib = _load()
    rgb = np.ascontiguousarray(rgb, dtype=np.uint8).ravel()
    model = rgb if model_rgb is None else np.ascontiguousarray(model_rgb, dtype=np.uint8).ravel()
    seeds = np.ascontiguousarray(seeds, dtype=np.uint8).ravel()
    n = width * height
    out = {
        "hist": np.zeros(96, np.int32), "counts": np.zeros(2, np.int32),
        "pos": np.zeros(n, np.int32), "neg": np.zeros(n, np.int32),
        "up": np.zeros(n, np.int32), "down": np.zeros(n, np.int32),
        "left": np.zeros(n, np.int32), "right": np.zeros(n, np.int32),
        "tpos_f": np.zeros(n, np.float64), "tneg_f": np.zeros(n, np.float64),
        "right_f": np.zeros(n, np.float64), "down_f": np.zeros(n, np.float64),
    }
    bsum, pairs, beta = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_double(0.0)
    rc = lib.oracle_dataterm_model(width, height, _ptr(rgb, ctypes.c_uint8), _ptr(model, ctypes.c_uint8),
                                   _ptr(seeds, ctypes.c_uint8), int(lam),
                             _ptr(out["hist"], ctypes.c_int32), _ptr(out["counts"], ctypes.c_int32),
                             ctypes.byref(bsum), ctypes.byref(pairs), ctypes.byref(beta),
                             *[_ptr(out[k], ctypes.c_int32) for k in ("pos", "neg", "up", "down", "left", "right")],
                             *[_ptr(out[k], ctypes.c_double) for k in ("tpos_f", "tneg_f", "right_f", "down_f")])
    if rc != 0:
        raise RuntimeError(f"oracle_dataterm failed: {rc}")
    out["hist"] = out["hist"].reshape(2, 3, 16)
    out["beta_sum"], out["pairs"], out["beta"] = int(bsum.value), int(pairs.value), float(beta.value)
    for k in ("pos", "neg", "up", "down", "left", "right", "tpos_f", "tneg_f", "right_f", "down_f"):
        out[k] = out[k].reshape(height, width)
    return out


def kde_gauss(sigma):
    """The 101 kernel values of the reference's K_InitGaussian (float32)."""
    g = np.zeros(101, np.float32)
    _load().oracle_kde_gauss(int(sigma), g.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    return g


def kde_select_colours(rgb, width, height, x, y, size=5):
    """Seed colours that selectOnMap(x, y) adds (DataTerm.cpp:111-123: x
    offset outer, y offset inner), pixels outside the image skipped."""
    img = np.asarray(rgb, np.uint8).reshape(height, width, 3)
    out = [img[yy, xx] for xx in range(x, x + size) for yy in range(y, y + size)
           if 0 <= xx < width and 0 <= yy < height]
    return np.array(out, np.uint8).reshape(-1, 3)


def dataterm_kde(width, height, rgb, seed_colours, sigma=5):
    """Kernel-density data term of the reference over the seed colours (N x 3
    u8, selection order). Returns dict with pos, neg (int32), p, tpos_f, tneg_f."""
    lib = _load()
    rgb = np.ascontiguousarray(rgb, dtype=np.uint8).ravel()
    sc = np.ascontiguousarray(seed_colours, dtype=np.uint8).reshape(-1, 3)
    n = width * height
    out = {"pos": np.zeros(n, np.int32), "neg": np.zeros(n, np.int32), "tpos_f": np.zeros(n, np.float64),
           "tneg_f": np.zeros(n, np.float64), "p": np.zeros(n, np.float64)}
    rc = lib.oracle_dataterm_kde(width, height, _ptr(rgb, ctypes.c_uint8), int(sc.shape[0]),
                                 _ptr(sc.ravel(), ctypes.c_uint8), int(sigma), _ptr(out["pos"], ctypes.c_int32),
                                 _ptr(out["neg"], ctypes.c_int32), _ptr(out["tpos_f"], ctypes.c_double),
                                 _ptr(out["tneg_f"], ctypes.c_double), _ptr(out["p"], ctypes.c_double))
    if rc != 0:
        raise RuntimeError(f"oracle_dataterm_kde failed: {rc}")
    return {k: v.reshape(height, width) for k, v in out.items()}


def refpr(width, height, pos, neg, edges=None, penalty=0, threads=0):
    """Reference-schedule push-relabel on the CPU. Returns (flow, label, iterations, printed_energy)."""
    lib = _load()
    pos, neg = _i32(pos).ravel(), _i32(neg).ravel()
    up, down, left, right = _edges(edges)
    label = np.zeros(width * height, np.int32)
    flow, printed = ctypes.c_int64(0), ctypes.c_int64(0)
    it = lib.oracle_refpr_solve(width, height, _ptr(pos, ctypes.c_int32), _ptr(neg, ctypes.c_int32),
                                _ptr(up, ctypes.c_int32), _ptr(down, ctypes.c_int32), _ptr(left, ctypes.c_int32),
                                _ptr(right, ctypes.c_int32), int(penalty), int(threads),
                                _ptr(label, ctypes.c_int32), ctypes.byref(flow), ctypes.byref(printed))
    if it < 0:
        raise RuntimeError(f"oracle_refpr_solve failed: {it}")
    return int(flow.value), label.reshape(height, width), int(it), int(printed.value)
