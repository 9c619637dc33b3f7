"""CPU-only checks of the product library: it loads, exports every symbol the
headers declare, and its host-only generators are deterministic. No GPU work."""
import os
import re
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _exports():
    import gcfg

    out = subprocess.run(["nm", "-D", "--defined-only", gcfg.LIB_PATH], check=True, stdout=subprocess.PIPE,
                         text=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_c_abi_exports_every_declared_function(g):
    hdr = open(os.path.join(ROOT, "include", "gcfg.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    declared = set(re.findall(r"\b(gcfg_\w+)\s*\(", hdr))
    assert len(declared) > 40
    missing = declared - _exports()
    assert not missing, sorted(missing)


def test_cpp_compat_exports(g):
    out = subprocess.run(["nm", "-D", "-C", "--defined-only", g.LIB_PATH], check=True, stdout=subprocess.PIPE,
                         text=True).stdout
    for hdr in ("GraphCut.h", "DataTerm.h"):
        text = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", hdr)).read(), flags=re.S)
        names = set(re.findall(r"\b(GC_\w+)\s*\(", text))
        if hdr == "DataTerm.h":
            names = {"DataTerm::" + n for n in re.findall(r"\b(?:int \*|void )(\w+)\s*\(", text)}
        assert names
        for n in names:
            assert re.search(r"\b" + re.escape(n) + r"\(", out), n


def test_default_options(g):
    o = g.default_options()
    assert o.sweeps_per_gr > 0 and o.inner_iters > 0 and o.steps_per_poll > 0


def test_generators_deterministic(g):
    a = g.gen_ellipse_image(50, 40, 3)
    b = g.gen_ellipse_image(50, 40, 3)
    c = g.gen_ellipse_image(50, 40, 4)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    # the ellipse interior is drawn around the foreground mean
    fg = a[20, 25].astype(int)
    assert abs(fg[0] - 200) < 90 and abs(fg[2] - 60) < 90
    s = g.gen_seed_map(100, 80)
    assert (s == 1).sum() == 400 and (s == 2).sum() == 1600
    pos, neg = g.gen_stress_tlinks(64, 64, 1)
    assert ((pos == 0) | (neg == 0)).all()
    v = pos + neg
    assert v.min() >= 32 and v.max() <= (1 << 20)
    w, h = g.gen_batch_shapes(200, 9, 128.0 ** 2, 1024.0 ** 2)
    area = w.astype(float) * h
    assert area.min() > 0.9 * 128 ** 2 and area.max() < 1.1 * 1024 ** 2
    assert (w / h).min() > 0.45 and (w / h).max() < 2.2
    p = g.gen_perlin_image(70, 30, 2)
    assert (p[..., 0] == p[..., 1]).all() and p.std() > 5


def test_library_fails_loudly_without_gpu_work(g):
    """No CPU fallback: a graph on a machine without a GPU errors instead of solving."""
    if g.gpu_available():
        return
    import pytest

    with pytest.raises(g.GcfgError):
        g.Graph(10, 10)


def test_tune_cache_roundtrip(g, tmp_path):
    path = str(tmp_path / "tune.txt")
    assert g.tune_lookup(path, "batch:256") is None
    o = g.default_options()
    o.inner_iters, o.sweeps_per_gr, o.gr_rounds = 2, 4, 8
    g.tune_store(path, "batch:256", o)
    o2 = g.default_options()
    o2.inner_iters = 16
    g.tune_store(path, "graph:4096x4096", o2)
    g.tune_store(path, "batch:256", o)  # replaces, does not duplicate
    got = g.tune_lookup(path, "batch:256")
    assert g.options_dict(got) == g.options_dict(o)
    assert g.tune_lookup(path, "graph:4096x4096").inner_iters == 16
    assert open(path).read().count("batch:256") == 1
