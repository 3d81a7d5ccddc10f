"""qs3x_ham_to_vec_wave_vector_noalloc against qs3x_ham_to_vec_wave_vector.

The no-allocation entry point runs the same H.v kernels (ham_to_vec_wave_vector,
ham2vec_v3.f90:37-72) with both host vectors staged through the library's
static per-device array.  Two calls in a row, the second on a smaller sector,
must each equal the allocating entry point bit for bit (the kernels are
deterministic) and the oracle within 1e-12; a sector beyond the staging limit
is refused with an error.
"""
import ctypes

import numpy as np
import pytest

from oracle import qs3 as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _q3():
    from ed_amd import qs3 as q3
    return q3


def _vec(dim, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1, 1, dim) + 1j * rng.uniform(-1, 1, dim)


def test_noalloc_hv_matches_allocating_entry_point_twice():
    q3 = _q3()
    # first call: 466,089 rows (L = 40, N_down = 7, k = 0, XXZ)
    print("[gpu] noalloc hv: L=40 N=7 k=0", flush=True)
    gm1 = q3.Model(40, 40, 1, 1, q3.chain_bonds(40, [(1, 1.0, 0.8)]))
    s1 = q3.Sector.build(gm1, 7, (0.0, 0.0, 0.0))
    assert s1.dim == 466089
    v1 = _vec(s1.dim, 1)
    ref1 = s1.ham_to_vec_wave_vector(v1)
    got1 = s1.ham_to_vec_wave_vector_noalloc(v1)
    assert np.array_equal(got1, ref1)
    # second call, smaller sector with complex phases (L = 20, N_down = 5, K = 3, J1-J2)
    print("[gpu] noalloc hv: L=20 N=5 K=3 J1-J2", flush=True)
    bonds = orc.chain_bonds(20, [(1, 1.0, 0.7), (2, 0.4, 0.5)])
    om = orc.Model(20, 20, 1, 1, bonds)
    rk = om.rk_from_K(3)
    osec = om.build_sector(5, rk)
    gm2 = q3.Model(20, 20, 1, 1, bonds)
    s2 = q3.Sector.build(gm2, 5, rk)
    assert s2.dim == osec.dim < s1.dim
    v2 = _vec(s2.dim, 2)
    got2 = s2.ham_to_vec_wave_vector_noalloc(v2)
    assert np.array_equal(got2, s2.ham_to_vec_wave_vector(v2))
    oref = om.hv(osec, v2)
    assert np.max(np.abs(got2 - oref)) <= 1e-12 * np.max(np.abs(oref))
    # the first sector again after the smaller call: nothing stale is read
    assert np.array_equal(s1.ham_to_vec_wave_vector_noalloc(v1), ref1)


def test_noalloc_hv_rejects_sectors_beyond_the_limit():
    q3 = _q3()
    print("[gpu] noalloc hv: limit", flush=True)
    gm = q3.Model(80, 80, 1, 1, q3.chain_bonds(80, [(1, 1.0, 1.0)]))
    s = q3.Sector.build(gm, 6, (0.0, 0.0, 0.0))
    assert s.dim > 2097152
    v = np.zeros(s.dim, dtype=np.complex128)
    out = np.zeros_like(v)
    rc = q3.lib().qs3x_ham_to_vec_wave_vector_noalloc(s._h, ctypes.c_void_p(out.ctypes.data),
                                                       ctypes.c_void_p(v.ctypes.data))
    assert rc != 0
    assert b"staging limit" in q3.lib().qs3x_last_error()
