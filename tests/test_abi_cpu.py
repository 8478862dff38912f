"""C-ABI checks that need no GPU: the library loads, exports every symbol the
public headers declare, the headers compile as C and C++, the Python
binding matches the headers, and the reference driver (main.cpp +
initialize.h) links against the library unchanged (link only)."""
import ctypes
import os
import re
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "molecular-dynamics_amd", "lib")
LIB = os.path.join(LIBDIR, "libtinymd_amd.so")
REF = "/root/reference"


def header_functions():
    names = []
    for h in ("md_abi.h", "tmd_ext.h"):
        text = open(os.path.join(INCLUDE, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b((?:md|tmd)_\w+)\s*\(", text)
    return sorted(set(names))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} not built")
    return ctypes.CDLL(LIB)


def test_header_declares_all_reference_entry_points():
    names = header_functions()
    md = [n for n in names if n.startswith("md_")]
    # anydsl_includes.h:10-48 declares 39 md_* functions
    assert len(md) == 39, md


def test_library_exports_every_declared_symbol(lib):
    out = subprocess.check_output(["nm", "-D", "--defined-only", LIB], text=True)
    exported = set(re.findall(r" T (\w+)", out))
    missing = [n for n in header_functions() if n not in exported]
    assert not missing, missing
    for n in header_functions():
        getattr(lib, n)


def test_python_binding_covers_headers():
    import sys
    sys.path.insert(0, os.path.join(ROOT, "molecular-dynamics_amd"))
    from tinymd_amd._lib import SIGNATURES
    assert set(header_functions()) == set(SIGNATURES), set(header_functions()) ^ set(SIGNATURES)


@pytest.mark.parametrize("lang,compiler", [("c", "gcc"), ("c++", "g++")])
def test_headers_compile(lang, compiler):
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "t." + ("c" if lang == "c" else "cpp"))
        with open(src, "w") as fh:
            fh.write('#include "md_abi.h"\n#include "tmd_ext.h"\nint main(void) { return 0; }\n')
        subprocess.check_call([compiler, "-std=c11" if lang == "c" else "-std=c++17", "-Wall", "-Werror",
                               "-fsyntax-only", "-I" + INCLUDE, src])


def test_headers_agree_with_reference_header():
    """One translation unit including the reference's ABI header and ours:
    any conflicting prototype is a compile error."""
    if not os.path.isdir(REF):
        pytest.skip("reference tree not present")
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "t.cpp")
        with open(src, "w") as fh:
            fh.write('#include "anydsl_includes.h"\n#include "md_abi.h"\nint main() { return 0; }\n')
        subprocess.check_call(["g++", "-std=c++17", "-fsyntax-only", "-I" + REF, "-I" + INCLUDE, src])


def test_reference_driver_links_against_library():
    """main.cpp and initialize.h from the reference tree, compiled and linked
    against libtinymd_amd.so unchanged (not run: no GPU here)."""
    if not os.path.isdir(REF):
        pytest.skip("reference tree not present")
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "md")
        subprocess.check_call(["g++", "-std=c++17", "-O0", "-I" + REF, os.path.join(REF, "main.cpp"), "-o", exe,
                               "-L" + LIBDIR, "-ltinymd_amd", "-Wl,-rpath," + LIBDIR,
                               "-Wl,-rpath-link,/opt/rocm/lib", "-Wl,--allow-shlib-undefined"])
        assert os.path.exists(exe)


def test_no_gpu_fails_loudly(lib):
    """Without a HIP device the library reports an error instead of falling
    back to anything."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "molecular-dynamics_amd"))
    import numpy as np
    import tinymd_amd as tm
    name, cus = tm.device_info()
    if cus > 0:
        pytest.skip("a GPU is present")
    m, x, v = tm.generate_fcc(4, 4, 4)
    with pytest.raises(tm.TinyMDError):
        tm.Simulation(m, x, v, tm.world_aabb(4, 4, 4), 2.8)
    tm.load().tmd_clear_error()


def test_host_lattice_generators_match_oracle():
    import sys
    sys.path.insert(0, os.path.join(ROOT, "molecular-dynamics_amd"))
    import numpy as np
    import pyoracle as po
    import tinymd_amd as tm
    for shape in ((4, 4, 4), (5, 3, 7), (8, 8, 8)):
        m, x, v = tm.generate_fcc(*shape)
        _, m2, x2, v2 = po.lattice(*shape)
        assert np.array_equal(x, x2) and np.array_equal(v, v2) and np.array_equal(m, m2)
    m, x, v = tm.generate_granular_gas(8, seed=3)
    ijk = np.rint(x / (1.6796 / 2)).astype(int)
    assert len(m) == (16 * 17 // 2) * 8  # i + j < 16 over a 16^3 vertex lattice, half the k parity
    assert np.all(ijk[:, 0] + ijk[:, 1] < 16) and np.all(ijk.sum(axis=1) % 2 == 0)
    assert np.all(np.abs(v) <= 1.0) and abs(v.mean()) < 0.05
    m64, _, _ = tm.generate_granular_gas(64)
    assert len(m64) == 528384


def test_replica_set_generator():
    """Counter-based replica parameters: deterministic, in range, and replica
    r's draw does not depend on how many replicas are asked for."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "molecular-dynamics_amd"))
    import numpy as np
    import tinymd_amd as tm
    c, r, t = tm.replica_set(seed=2024, nrep=64)
    assert c.min() >= 6 and c.max() <= 24
    assert np.all((r >= 0.70) & (r < 0.90)) and np.all((t >= 0.7) & (t < 2.0))
    c5, r5, t5 = tm.replica_set(seed=2024, nrep=5)
    assert np.array_equal(c[:5], c5) and np.array_equal(r[:5], r5) and np.array_equal(t[:5], t5)
    c2, _, _ = tm.replica_set(seed=2025, nrep=64)
    assert not np.array_equal(c, c2)
    # the 64-replica workload: sum of 4 nc^3 atoms
    assert int((4 * c.astype(np.int64) ** 3).sum()) == 1334984
    boxes = tm.replica_boxes(seed=2024, nrep=3)
    for (aabb, x, v), nc, rho, temp in zip(boxes, c, r, t):
        assert len(x) == 4 * nc ** 3
        assert aabb[1] == pytest.approx(nc * (4.0 / rho) ** (1.0 / 3.0))
        assert np.abs(v.mean(axis=0)).max() < 1e-12
        assert np.sum(v * v) / (3 * len(v) - 3) == pytest.approx(temp, rel=1e-12)


def test_kahan_finalisation():
    """The compensated FP32 sum of the FP32-Kahan pair kernels (the library's
    own inline helpers, evaluated on the host).  Kahan's bound: the error of
    the compensated sum is at most (2u + O(n u^2)) sum |x_i| with u = 2^-24,
    whatever n; the plain FP32 running sum has no such bound and is worse on
    average.  The exact sum of the FP32 inputs is math.fsum."""
    import math
    import sys
    sys.path.insert(0, os.path.join(ROOT, "molecular-dynamics_amd"))
    import numpy as np
    import tinymd_amd as tm
    L = tm.load()
    rng = np.random.default_rng(5)
    u = 2.0 ** -24
    err_k, err_n, err_wrong = [], [], []
    for trial in range(200):
        n = 80
        # pair-force-like terms: mixed signs and magnitudes, large cancellation
        x = (rng.standard_normal(n) * np.exp(rng.uniform(-3, 4, n))).astype(np.float32)
        exact = math.fsum(float(t) for t in x)
        got = L.tmd_kahan_sum(x.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n)
        bound = (2 * u + 4 * n * u * u) * float(np.abs(x.astype(np.float64)).sum())
        assert abs(got - exact) <= bound, (trial, got, exact, bound)
        err_k.append(abs(got - exact))
        s = np.float32(0.0)
        c = np.float32(0.0)
        for t in x:  # negative control: the finalisation with the sign flipped
            y = np.float32(t - c)
            tt = np.float32(s + y)
            c = np.float32(np.float32(tt - s) - y)
            s = tt
        err_wrong.append(abs((float(s) + float(c)) - exact))
        err_n.append(abs(float(np.cumsum(x, dtype=np.float32)[-1]) - exact))
    assert np.mean(err_k) < 0.25 * np.mean(err_n), (np.mean(err_k), np.mean(err_n))
    # the flipped sign adds the rounding error a second time instead of
    # removing it: about twice the compensated error
    assert np.mean(err_wrong) > 1.5 * np.mean(err_k), (np.mean(err_wrong), np.mean(err_k))


def test_report_time_format(capfd):
    """md_report_time prints the reference's header and the six values with
    a trailing blank (core/runtime.impala:318-337; the layout of the
    results logs, e.g. results/single_gpu_v100.txt:96-97)."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "molecular-dynamics_amd"))
    import tinymd_amd as tm
    L = tm.load()
    capfd.readouterr()
    L.md_report_time(154.0, 84.0829, 51.3265, 19.8786, 0.0, 0.296124)
    out = capfd.readouterr().out
    assert out == "Total Force Neigh Comm Balance Other\n154 84.0829 51.3265 19.8786 0 0.296124 \n", repr(out)
