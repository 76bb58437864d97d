"""The triangulation's exact predicates (surface-remesher_amd/csrc/dt_predicates.h),
compiled for the host with g++, against exact rational arithmetic on random,
near-degenerate and exactly degenerate records; the range limit is reported.
The same header runs on the device (tests/test_gpu_delaunay.py)."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from dt_cases import (degenerate_records, exact_incircle, exact_orient, records, sos_incircle,
                      sos_orient)

ROOT = Path(__file__).resolve().parent.parent
HARNESS = r'''
#include "dt_predicates.h"
#include <cstdio>
#include <vector>
int main(int argc, char **argv) {
    FILE *f = std::fopen(argv[1], "rb");
    std::vector<double> q;
    double v;
    while (std::fread(&v, sizeof v, 1, f) == 1) q.push_back(v);
    std::fclose(f);
    const size_t n = q.size() / 8;
    for (size_t i = 0; i < n; ++i) {
        const double *x = &q[8 * i];
        int err = 0;
        const int o = cvt::dt::orient2d(x[0], x[1], x[2], x[3], x[4], x[5], &err);
        const int c = cvt::dt::incircle(x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7], &err);
        // indices for the perturbation: a permutation of 10, 20, 30, 40 by record
        static const int perms[6][4] = {{10, 20, 30, 40}, {40, 30, 20, 10}, {20, 40, 10, 30},
                                        {30, 10, 40, 20}, {10, 40, 20, 30}, {40, 20, 30, 10}};
        const int *id = perms[i % 6];
        const int so = cvt::dt::orient2d_sos(x[0], x[1], x[2], x[3], x[4], x[5], id[0], id[1],
                                             id[2], &err);
        const int sc = cvt::dt::incircle_sos(x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7], id[0],
                                             id[1], id[2], id[3], &err);
        std::printf("%d %d %d %d %d\n", o, c, err, so, sc);
    }
    return 0;
}
'''


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ not found")
    d = tmp_path_factory.mktemp("dtpred")
    (d / "h.cpp").write_text(HARNESS)
    exe = d / "h"
    subprocess.run([gxx, "-O2", "-std=c++17", "-ffp-contract=off", "-I",
                    str(ROOT / "surface-remesher_amd" / "csrc"), str(d / "h.cpp"), "-o", str(exe)],
                   check=True, capture_output=True, timeout=300)

    def run(q):
        p = d / "q.bin"
        np.ascontiguousarray(q, np.float64).tofile(p)
        out = subprocess.run([str(exe), str(p)], capture_output=True, text=True, check=True,
                             timeout=120).stdout.split()
        return np.array(out, dtype=np.int64).reshape(-1, 5)
    return run


def test_exact_signs(harness):
    q = records()
    got = harness(q)
    assert (got[:, 2] == 0).all()
    want_o = np.array([exact_orient(r) for r in q])
    want_c = np.array([exact_incircle(r) for r in q])
    assert np.array_equal(got[:, 0], want_o)
    assert np.array_equal(got[:, 1], want_c)
    # the degenerate records are really there
    assert (want_o == 0).sum() >= 60 and (want_c == 0).sum() >= 80


def test_sign_conventions(harness):
    q = np.array([[0, 0, 1, 0, 0, 1, 0.25, 0.25],       # ccw; d inside
                  [0, 0, 0, 1, 1, 0, 0.25, 0.25],       # cw
                  [0, 0, 1, 0, 0, 1, 2.0, 2.0]],        # d outside
                 np.float64)
    got = harness(q)
    assert got[:, 0].tolist() == [1, -1, 1]
    assert got[:, 1].tolist() == [1, -1, -1]


def test_range_limit_reported(harness):
    # the exact stage cannot scale 1e-30 and 1e30 to one integer grid: the
    # filter cannot decide a collinear triple there, so the range error is set
    q = np.array([[1e-30, 1e-30, 1e30, 1e30, 2e30, 2e30, 0.0, 0.0]])
    got = harness(q)
    assert got[0, 2] & 2


PERMS = [(10, 20, 30, 40), (40, 30, 20, 10), (20, 40, 10, 30), (30, 10, 40, 20),
         (10, 40, 20, 30), (40, 20, 30, 10)]


def test_simulation_of_simplicity(harness):
    """where the exact sign is zero, the perturbed sign equals the leading
    term of the determinant expanded symbolically in eps (tests/dt_cases.py),
    for six index orders; elsewhere it equals the exact sign"""
    q = np.concatenate([degenerate_records(), records(seed=3, n_random=40)])
    got = harness(q)
    assert (got[:, 2] == 0).all()
    n_deg = 0
    for i, r in enumerate(q):
        idx = PERMS[i % 6]
        o, c = exact_orient(r), exact_incircle(r)
        assert got[i, 3] == (o if o else sos_orient(r, idx[:3])), (i, r)
        assert got[i, 4] == (c if c else sos_incircle(r, idx)), (i, r)
        assert got[i, 3] != 0 and got[i, 4] != 0
        n_deg += (o == 0) + (c == 0)
    assert n_deg >= 150
