"""CPU-side checks of the C ABI and host logic (no GPU compute).

The engine library must load and export every entry point include/*.h
declares; argument validation that happens before any device call is
exercised here; the workload generators are deterministic.
"""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(cvt_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    for n in ["cvt_create", "cvt_destroy", "cvt_set_density", "cvt_set_sites", "cvt_voronoi",
              "cvt_moments", "cvt_lloyd", "cvt_energy", "cvt_get_labels"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    import cvtamd
    L = C.CDLL(str(cvtamd.LIB_PATH))
    missing = [n for n in declared_symbols() if not hasattr(L, n)]
    assert not missing, missing


def test_reference_signature_shims_exported():
    # C++ linkage, the reference's exact parameter lists (include/gcvt_compat.h)
    import cvtamd
    L = C.CDLL(str(cvtamd.LIB_PATH))
    assert hasattr(L, "_Z4gCVTPsPfPbiii")                  # gCVT(short*, float*, bool*, int, int, int)
    assert hasattr(L, "_Z16discretization_dPdS_iPiiPfdi")  # discretization_d(double*, double*, int, int*, int, float*, double, int)


def test_python_binding_covers_header():
    from cvtamd import _lib
    assert declared_symbols() <= set(_lib.SIGNATURES), declared_symbols() - set(_lib.SIGNATURES)


def test_version_and_argument_errors_without_gpu():
    import cvtamd
    L = cvtamd.lib()
    assert L.cvt_version() == 100
    h = C.c_void_p()
    assert L.cvt_create(C.byref(h), 0, 64, 10) == -1
    assert b"grid" in L.cvt_last_error()
    assert L.cvt_create(C.byref(h), 64, 64, -5) == -1
    assert L.cvt_create(None, 64, 64, 5) == -1
    assert L.cvt_lloyd(None, 1, 1.0) == -1
    assert L.cvt_destroy(None) == 0


def test_python_error_type():
    import cvtamd
    with pytest.raises(cvtamd.CvtError) as e:
        cvtamd.Engine(0, 0, 1)
    assert e.value.status == -1


def test_workloads_deterministic():
    from cvtamd import workloads as wl
    a = wl.config2(5, n=256, num_sites=300)
    b = wl.config2(5, n=256, num_sites=300)
    assert np.array_equal(a.density, b.density) and np.array_equal(a.sites, b.sites)
    c = wl.config2(6, n=256, num_sites=300)
    assert not np.array_equal(a.sites, c.sites)


def test_workload_shapes_and_domains():
    from cvtamd import workloads as wl
    w1 = wl.config1()
    assert w1.n == 512 and w1.num_sites == 1024 and w1.iterations == 50
    assert (w1.density == 1).all() and w1.sites.dtype == np.float32
    assert (w1.sites >= 0).all() and (w1.sites < 512).all()
    w2 = wl.config2(2, n=256, num_sites=500)
    m = wl.disk_mask(256)
    assert (w2.density[~m] == 0).all() and (w2.density[m] >= 0.05).all()
    px = np.floor(w2.sites).astype(int)
    assert (w2.density[px[:, 1], px[:, 0]] > 0).all()
    w3 = wl.config3(3, n=256, num_sites=400)
    assert (w3.density > 0).sum() > 0.3 * 256 * 256
    assert (w3.density >= 0).all()


def test_batch_patch_rules():
    from cvtamd import workloads as wl
    for s in range(6):
        p = wl.batch_patch(s)
        assert p.n % 64 == 0 and 256 <= p.n <= 2048
        assert 256 <= p.num_sites <= 16384
        assert 4 <= p.meta["components"] <= 12


def test_c_host_compiles_links_and_runs(tmp_path):
    """include/cvt.h is plain C: the documented call sequence compiles with
    gcc, links to the library and runs (an error report without a GPU)."""
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not found")
    from conftest import ROOT
    lib = ROOT / "surface-remesher_amd" / "lib"
    exe = tmp_path / "abi_example"
    subprocess.run([gcc, "-std=c99", "-Wall", "-Werror", "-I", str(ROOT / "include"),
                    str(ROOT / "tests" / "c" / "abi_example.c"), "-L", str(lib), "-lcvt_hip",
                    f"-Wl,-rpath,{lib}", "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("version ")


# ------------------------------------------------- reference site initialiser --
def test_ref_init_sites_kat_and_oracle(orc):
    """cvt_ref_init_sites runs on the host: the zero-state KATs (first site
    (0, 437) at 512 and (0, 1748) at 2048 for a uniform density) and the
    oracle's putConstrains + randomPoints bit for bit, with constraints, a
    masked-out disk and a continued generator state."""
    from cvtamd import compat
    from cvtamd import workloads as wl
    for n, first in ((512, (0, 437)), (2048, (0, 1748))):
        m, d = compat.ref_init_sites(np.ones((n, n), np.float32), 1)
        ys, xs = np.nonzero(m[:, :, 0] != -32768)
        assert list(zip(xs.tolist(), ys.tolist())) == [first] and d == 3
        assert m[first[1], first[0]].tolist() == list(first)
    for n, num, seed in ((64, 40, 1), (300, 900, 2), (512, 1000, 3)):
        g = wl.rng(seed)
        rho = wl.gaussian_mixture(n, g, 5, 0.05, wl.disk_mask(n))
        mask = np.zeros((n, n), np.uint8)
        ys, xs = np.nonzero(rho > 0)
        pick = g.integers(0, len(xs), 7)
        mask[ys[pick], xs[pick]] = 1
        ref = orc.ref_put_constraints(mask)
        kiss = orc.kiss(0)
        orc.ref_random_points(ref, rho, num, kiss)
        st = np.zeros(4, np.uint32)
        got, _ = compat.ref_init_sites(rho, num, mask, rng_state=st)
        assert np.array_equal(got, ref)
        assert st.tolist() == [kiss.z, kiss.w, kiss.jsr, kiss.jcong]
        # a second call continues the sequence, as the reference's globals do
        orc.ref_random_points(ref, rho, 10, kiss)
        again, _ = compat.ref_init_sites(rho, num + 10, mask)
        assert np.array_equal(again, ref)


def test_ref_init_sites_refuses_endless_loops():
    import cvtamd
    from cvtamd import compat
    n = 16
    rho = np.zeros((n, n), np.float32)
    rho[3, 4] = 1.0
    rho[5, 6] = 2.0
    m, _ = compat.ref_init_sites(rho, 2)
    assert sorted(zip(*np.nonzero(m[:, :, 0] != -32768))) == [(3, 4), (5, 6)]
    with pytest.raises(cvtamd.CvtError):          # three sites, two pixels
        compat.ref_init_sites(rho, 3)
    mask = np.zeros((n, n), np.uint8)
    mask[3, 4] = 1
    with pytest.raises(cvtamd.CvtError):          # the masked pixel is taken
        compat.ref_init_sites(rho, 2, mask)
    with pytest.raises(cvtamd.CvtError):
        compat.ref_init_sites(np.full((n, n), np.nan, np.float32), 1)
    m, d = compat.ref_init_sites(np.zeros((n, n), np.float32), 0)
    assert (m == -32768).all() and d == 0


def test_cpp_caller_builds_against_shim_header(tmp_path):
    """A C++ host that includes include/gcvt_compat.h (the reference's two
    signatures, C++ linkage) and include/cvt.h compiles with g++, links the
    mangled gCVT and discretization_d from libcvt_hip.so, and runs the host
    part of centroidalVoronoi (the site initialiser) without a GPU."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    lib = ROOT / "surface-remesher_amd" / "lib"
    exe = tmp_path / "gcvt_caller"
    subprocess.run([gxx, "-std=c++17", "-Wall", "-Werror", "-I", str(ROOT / "include"),
                    str(ROOT / "tests" / "c" / "gcvt_caller.cpp"), "-L", str(lib), "-lcvt_hip",
                    f"-Wl,-rpath,{lib}", "-o", str(exe)], check=True)
    out = subprocess.run([str(exe), "host"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "first site 0 437" in out.stdout
