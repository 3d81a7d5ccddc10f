"""CPU-side checks of libqs3x: it loads without a GPU, exports every entry
point include/qs3x.h declares, and its host-only pieces (small eigensolvers,
continued fraction, start-vector formula) are right."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "ed_amd", "libqs3x.so")
HDR = os.path.join(ROOT, "include", "qs3x.h")


def _declared():
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(qs3x_[a-z0-9_]+)\s*\(", txt)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("ed_amd/libqs3x.so is not built (run ed_amd/build.py)")
    return ctypes.CDLL(LIB)


def test_exports_every_declared_symbol(lib):
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n


def test_python_mirror_imports():
    from ed_amd import qs3 as q3
    assert q3.wavevector(1, 4) == pytest.approx(np.pi / 2)
    # the exact expression of read_ip: 2 * acos(-1) * K / L
    assert q3.wavevector(3, 7) == 2.0 * np.arccos(-1.0) * 3.0 / 7.0


def test_host_tridiagonal_solver(lib):
    lib.qs3x_test_tridiag_eigenvalue.restype = ctypes.c_double
    lib.qs3x_test_tridiag_eigenvalue.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    lib.qs3x_test_tridiag_vector.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                             ctypes.c_void_p]
    rng = np.random.default_rng(0)
    for n in (1, 2, 3, 10, 200, 1500):
        d = rng.normal(size=n) * 3
        e = np.abs(rng.normal(size=max(n - 1, 1)))
        T = np.diag(d) + (np.diag(e[: n - 1], 1) + np.diag(e[: n - 1], -1) if n > 1 else 0)
        w, V = np.linalg.eigh(T)
        for k in (0, 1, n - 1):
            if k >= n:
                continue
            got = lib.qs3x_test_tridiag_eigenvalue(d.ctypes.data, e.ctypes.data, n, k)
            assert got == pytest.approx(w[k], abs=1e-12 * max(1.0, np.max(np.abs(w))))
        z = np.zeros(n)
        lib.qs3x_test_tridiag_vector(d.ctypes.data, e.ctypes.data, n, ctypes.c_double(w[0]), z.ctypes.data)
        assert abs(abs(np.dot(z, V[:, 0])) - 1.0) < 1e-9


def test_host_dense_symmetric_solver(lib):
    lib.qs3x_test_sym_eig.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    rng = np.random.default_rng(1)
    for n in (1, 4, 17, 60):
        A = rng.normal(size=(n, n))
        A = A + A.T
        ref = np.linalg.eigvalsh(A)
        a = A.copy()
        w = np.zeros(n)
        V = np.zeros((n, n))
        lib.qs3x_test_sym_eig(n, a.ctypes.data, w.ctypes.data, V.ctypes.data)
        assert np.allclose(w, ref, atol=1e-12 * max(1, np.max(np.abs(ref))))
        assert np.allclose(A @ V, V * w[None, :], atol=1e-10)


def test_continued_fraction_matches_oracle():
    from ed_amd import qs3 as q3
    from oracle import qs3 as orc
    rng = np.random.default_rng(2)
    a = rng.normal(size=50)
    b = np.concatenate([[0.0], np.abs(rng.normal(size=49))])
    om = np.linspace(-2, 5, 1001)
    for quirk in (True, False):
        s1 = q3.dsf_analysis(a, b, 1.7, -0.3, om, 0.04, reference_quirk=quirk)
        s2 = orc.continued_fraction(a, b, 1.7, -0.3, om, 0.04, reference_quirk=quirk)
        assert np.allclose(s1, s2, rtol=1e-13, atol=0)


def test_start_vector_formula_is_deterministic():
    from ed_amd import qs3 as q3
    a = q3.start_vector(1000, 42)
    b = q3.start_vector(500, 42, row0=500)
    assert np.array_equal(a[500:], b)
    assert np.all(np.abs(a.real) < 1) and np.all(np.abs(a.imag) < 1)
    assert abs(np.mean(a.real)) < 0.1 and abs(np.mean(a.imag)) < 0.1


def test_arrowhead_lowest_pairs(lib):
    """Structured eigensolver of the thick-restart projection (locked Ritz
    values w coupled to row k, tridiagonal tail) against numpy's eigh, including
    a pure tridiagonal (k = 0), a degenerate locked block and a large size."""
    lib.qs3x_test_arrow_pairs.restype = None
    lib.qs3x_test_arrow_pairs.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 4 + [ctypes.c_int] + \
        [ctypes.c_void_p] * 2
    rng = np.random.default_rng(5)
    for k, m, nv, degen in ((0, 40, 4, False), (4, 60, 4, False), (8, 300, 8, False), (4, 50, 4, True),
                            (2, 3, 2, False), (0, 1, 1, False), (1, 2, 1, False)):
        w = np.sort(rng.normal(size=k)) if k else np.zeros(1)
        if degen and k >= 2:
            w[1] = w[0]
        c = rng.normal(size=max(k, 1))
        a = rng.normal(size=m - k) if m > k else np.zeros(1)
        b = rng.uniform(0.1, 1.0, size=max(m - k - 1, 1))
        A = np.zeros((m, m))
        for j in range(k):
            A[j, j] = w[j]
            A[j, k] = A[k, j] = c[j]
        for i in range(k, m):
            A[i, i] = a[i - k]
            if i + 1 < m:
                A[i, i + 1] = A[i + 1, i] = b[i - k]
        lam = np.zeros(nv)
        Z = np.zeros((m, nv))
        lib.qs3x_test_arrow_pairs(k, m, w.ctypes.data, c.ctypes.data, a.ctypes.data, b.ctypes.data, nv,
                                  lam.ctypes.data, Z.ctypes.data)
        ev = np.linalg.eigvalsh(A)
        assert np.allclose(lam, ev[:nv], atol=1e-12 * max(1, np.max(np.abs(ev)))), (k, m, lam, ev[:nv])
        res = A @ Z - Z * lam
        assert np.max(np.abs(res)) < 1e-9, (k, m, np.max(np.abs(res)))
        assert np.allclose(Z.T @ Z, np.eye(nv), atol=1e-9)


def test_env_options_are_validated():
    """QS3X_OPTIONS is parsed before the first model is created (no GPU needed
    for that step): an unknown name or a bad value is reported with the
    variable's name; valid settings pass on to the device set-up."""
    import subprocess
    import sys
    code = (
        "import ctypes, numpy as np\n"
        "L = ctypes.CDLL('ed_amd/libqs3x.so')\n"
        "L.qs3x_last_error.restype = ctypes.c_char_p\n"
        "p = np.array([1, 2, 2, 3, 3, 1], dtype=np.int32); j = np.ones(3)\n"
        "h = ctypes.c_void_p()\n"
        "rc = L.qs3x_model_create(3, 3, 1, 1, 3, p.ctypes.data_as(ctypes.c_void_p), "
        "j.ctypes.data_as(ctypes.c_void_p), j.ctypes.data_as(ctypes.c_void_p), ctypes.byref(h))\n"
        "print(rc, L.qs3x_last_error().decode() if rc else '')\n")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    for env, expect in (("hv_kernel=7", "QS3X_OPTIONS: hv_kernel must be 0..3"),
                        ("nonsense=1", "QS3X_OPTIONS: qs3x_set_option: unknown option nonsense"),
                        ("tr_reorth", "QS3X_OPTIONS: expected name=value"),
                        ("tr_reorth=x", "QS3X_OPTIONS: bad value"),
                        ("hv_chunk=25", "QS3X_OPTIONS: hv_chunk must be -1..24"),
                        ("hv_xcd=65", "QS3X_OPTIONS: hv_xcd must be 0..64")):
        out = subprocess.run([sys.executable, "-c", code], cwd=root, capture_output=True, text=True, timeout=120,
                             env=dict(os.environ, QS3X_OPTIONS=env, HIP_VISIBLE_DEVICES=""))
        assert expect in out.stdout, (env, out.stdout, out.stderr)
    out = subprocess.run([sys.executable, "-c", code], cwd=root, capture_output=True, text=True, timeout=120,
                         env=dict(os.environ, QS3X_OPTIONS="tr_check_every=5,tr_reorth=2,tr_cgs_passes=1,hv_chunk=3",
                                  HIP_VISIBLE_DEVICES=""))
    assert "QS3X_OPTIONS" not in out.stdout, out.stdout
