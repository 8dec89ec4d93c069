"""The 3x3 expm / logm / sqrtm / Frechet routines of the HIP library
(csrc/tsx_mat33.h), compiled for the host with g++, against scipy.linalg.

These are the routines the FIRE cell-filter kernels run per system
(reference: torch_sim/math.py:106-189 expm_frechet, 614-745 matrix_log_33,
torch.matrix_exp in optimizers/fire.py:424 and cell_filters.py:333).
"""

import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest
import scipy.linalg as sla

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "torch-sim_amd", "csrc")


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("mat33") / "libmat33.so"
    src = os.path.join(ROOT, "tests", "native", "mat33_capi.cpp")
    subprocess.run([gxx, "-O2", "-std=c++17", "-shared", "-fPIC", "-I", CSRC, src, "-o", str(out)],
                   check=True)
    so = ctypes.CDLL(str(out))
    for name in ("c_expm", "c_logm", "c_sqrtm"):
        getattr(so, name).argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    so.c_expm_frechet.argtypes = [ctypes.c_void_p] * 4
    return so


def _call(fn, *mats):
    ins = [np.ascontiguousarray(m, dtype=np.float64) for m in mats]
    outs = [np.zeros((3, 3)) for _ in range(2 if fn.__name__ == "c_expm_frechet" else 1)]
    fn(*[m.ctypes.data for m in ins], *[o.ctypes.data for o in outs])
    return outs if len(outs) > 1 else outs[0]


def _cases(seed, n, scale):
    rng = np.random.default_rng(seed)
    return [rng.normal(scale=scale, size=(3, 3)) for _ in range(n)]


@pytest.mark.parametrize("scale", [1e-3, 0.1, 0.7, 2.0])
def test_expm_matches_scipy(lib, scale):
    for a in _cases(1, 20, scale):
        ref = sla.expm(a)
        got = _call(lib.c_expm, a)
        # ||A||_1 up to ~6 at scale 2 (no squaring below theta_13 = 5.37): the
        # Pade denominator solve then costs a few hundred ulps
        tol = 1e-13 if scale < 1.0 else 2e-12
        assert np.abs(got - ref).max() <= tol * max(1.0, np.abs(ref).max())


def test_expm_special_cases(lib):
    assert np.array_equal(_call(lib.c_expm, np.zeros((3, 3))), np.eye(3))
    d = np.diag([0.3, -1.2, 2.5])
    assert np.allclose(_call(lib.c_expm, d), np.diag(np.exp([0.3, -1.2, 2.5])), rtol=1e-14)
    # nilpotent: exp(N) = I + N + N^2/2 exactly
    n = np.array([[0, 1.5, -2.0], [0, 0, 0.25], [0, 0, 0]])
    assert np.allclose(_call(lib.c_expm, n), np.eye(3) + n + n @ n / 2, rtol=0, atol=1e-15)
    # rotation generator: exp of a skew matrix is orthogonal
    k = np.array([[0, -0.4, 0.2], [0.4, 0, -0.9], [-0.2, 0.9, 0]])
    r = _call(lib.c_expm, k)
    assert np.allclose(r @ r.T, np.eye(3), atol=1e-14)


@pytest.mark.parametrize("scale", [1e-3, 0.1, 0.5, 2.0])
def test_expm_frechet_matches_scipy(lib, scale):
    for a, e in zip(_cases(2, 20, scale), _cases(3, 20, 1.0)):
        ea_ref, l_ref = sla.expm_frechet(a, e)
        ea, l = _call(lib.c_expm_frechet, a, e)
        assert np.abs(ea - ea_ref).max() <= 1e-13 * max(1.0, np.abs(ea_ref).max())
        assert np.abs(l - l_ref).max() <= 1e-12 * max(1.0, np.abs(l_ref).max())


def test_expm_frechet_finite_difference(lib):
    for a, e in zip(_cases(4, 5, 0.5), _cases(5, 5, 1.0)):
        h = 1e-6
        fd = (sla.expm(a + h * e) - sla.expm(a - h * e)) / (2 * h)
        _, l = _call(lib.c_expm_frechet, a, e)
        assert np.abs(l - fd).max() < 1e-8 * max(1.0, np.abs(fd).max())


def test_expm_frechet_commuting_case(lib):
    # for commuting A, E: L(A, E) = E exp(A)
    a = np.diag([0.2, -0.5, 1.1])
    e = np.diag([1.0, 2.0, -3.0])
    _, l = _call(lib.c_expm_frechet, a, e)
    assert np.allclose(l, e @ sla.expm(a), rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("scale", [1e-4, 0.05, 0.3, 0.8])
def test_logm_matches_scipy(lib, scale):
    for b in _cases(6, 20, scale):
        a = sla.expm(b)  # principal branch guaranteed for small b
        ref = np.real(sla.logm(a))
        got = _call(lib.c_logm, a)
        assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


def test_logm_special_cases(lib):
    assert np.abs(_call(lib.c_logm, np.eye(3))).max() == 0.0
    d = np.diag([0.5, 2.0, 7.0])
    assert np.allclose(_call(lib.c_logm, d), np.diag(np.log([0.5, 2.0, 7.0])), rtol=1e-14, atol=1e-15)
    # symmetric positive definite: log = V log(L) V^T
    rng = np.random.default_rng(7)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    s = q @ np.diag([0.3, 1.0, 4.0]) @ q.T
    assert np.allclose(_call(lib.c_logm, s), q @ np.diag(np.log([0.3, 1.0, 4.0])) @ q.T, atol=1e-13)


def test_sqrtm_squares_back(lib):
    for b in _cases(8, 10, 0.6):
        a = sla.expm(b)
        r = _call(lib.c_sqrtm, a)
        assert np.abs(r @ r - a).max() < 1e-13 * np.abs(a).max()
        assert np.allclose(r, sla.sqrtm(a).real, atol=1e-12)
