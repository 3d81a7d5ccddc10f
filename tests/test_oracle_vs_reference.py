"""Pins the C oracle against the reference's own Fortran modules.

oracle/_ref/libqs3ref.so is built by oracle/build_ref.sh from the reference
sources (state_lists.f90, ham2vec_v3.f90, input_param.f90,
get_expectation_values.f90) and a harness that sets their module variables.
Integers and sector lists must agree bit for bit; H.v and observables to
roundoff.  Skipped where the reference build is absent (it is never needed
on a GPU box)."""
import numpy as np
import pytest

from oracle import qs3 as orc

pytestmark = pytest.mark.skipif(not orc.RefLib.available(), reason="oracle/_ref not built")

CHAINS = [
    (12, 4, 0, [(1, 1.0, 1.0)]),
    (12, 4, 3, [(1, 1.0, 0.7), (2, 0.5, 0.5)]),
    (10, 5, 5, [(1, 1.0, 1.0)]),
    (14, 7, 2, [(1, 1.0, 0.8)]),
    (18, 3, 1, [(1, 1.0, 1.0), (3, 0.3, -0.2)]),
    (16, 8, 0, [(1, 1.0, 1.0)]),
]

TORI = [
    (6, 6, 3, (0, 0, 0)),
    (6, 6, 3, (1, 2, 0)),
    (4, 4, 4, (1, 0, 0)),
    (4, 3, 3, (2, 1, 0)),
]


def _ref_setup(NOS, LX, LY, LZ, nod, K, bonds):
    ref = orc.RefLib()
    ref.setup(NOS, LX, LY, LZ, nod, K, bonds)
    return ref


@pytest.mark.parametrize("L,N,K,fam", CHAINS)
def test_chain_sector_and_hv(L, N, K, fam):
    bonds = orc.chain_bonds(L, fam)
    m = orc.Model(L, L, 1, 1, bonds)
    rk = m.rk_from_K(K)
    sec = m.build_sector(N, rk)
    fast = m.build_sector_chain(N, rk)
    ref = _ref_setup(L, L, 1, 1, N, (K, 0, 0), bonds)
    D = ref.build_sector()
    ls, lr = ref.lists()
    assert D == sec.dim == fast.dim
    assert np.array_equal(ls, sec.list_s) and np.array_equal(ls, fast.list_s)
    assert np.array_equal(lr, sec.list_r) and np.array_equal(lr, fast.list_r)
    assert np.array_equal(ref.explist(), sec.explist)
    rng = np.random.default_rng(1)
    v = rng.uniform(-1, 1, D) + 1j * rng.uniform(-1, 1, D)
    a, b = ref.hv(v), m.hv(sec, v)
    assert np.max(np.abs(a - b)) <= 1e-14 * np.max(np.abs(a))


@pytest.mark.parametrize("LX,LY,N,K", TORI)
def test_torus_sector_and_hv(LX, LY, N, K):
    NOS = LX * LY
    bonds = orc.square_bonds(LX, LY, 1.0)
    m = orc.Model(NOS, LX, LY, 1, bonds)
    rk = (orc.wavevector(K[0], LX), orc.wavevector(K[1], LY), 0.0)
    sec = m.build_sector(N, rk)
    ref = _ref_setup(NOS, LX, LY, 1, N, K, bonds)
    D = ref.build_sector()
    ls, lr = ref.lists()
    assert D == sec.dim
    assert np.array_equal(ls, sec.list_s)
    assert np.array_equal(lr, sec.list_r)
    rng = np.random.default_rng(2)
    v = rng.uniform(-1, 1, D) + 1j * rng.uniform(-1, 1, D)
    assert np.max(np.abs(ref.hv(v) - m.hv(sec, v))) <= 1e-14 * np.max(np.abs(ref.hv(v)))


def test_primitives_against_reference():
    """list_fly, j_flip_ni, representative_SQ (rank and ell), checkstate_SQ."""
    L, N = 11, 4
    bonds = orc.chain_bonds(L, [(1, 1.0, 1.0)])
    m = orc.Model(L, L, 1, 1, bonds)
    ref = _ref_setup(L, L, 1, 1, N, (3, 0, 0), bonds)
    rk = m.rk_from_K(3)
    THS = m.comb(L, N)
    for t in range(1, THS + 1):
        n = m.unrank(t, N)
        assert np.array_equal(n, ref.unrank(t, N))
        assert m.rank(n) == t
        r1 = m.representative(n)
        r2 = ref.representative(n)
        assert r1[0] == r2[0] and np.array_equal(r1[1], r2[1]) and np.array_equal(r1[2], r2[2])
        assert m.checkstate(t, N, rk)[0] == ref.checkstate(t, N)
    # j_flip on every antiparallel pair of a few configurations
    for t in (1, 17, 100, THS):
        n = m.unrank(t, N)
        s = set(n.tolist())
        for i in range(1, L + 1):
            for j in range(i + 1, L + 1):
                if (i in s) != (j in s):
                    assert np.array_equal(m.j_flip(i, j, n), ref.j_flip(i, j, n))


def test_torus_representative_ell():
    LX, LY, N = 4, 3, 3
    NOS = LX * LY
    bonds = orc.square_bonds(LX, LY, 1.0)
    m = orc.Model(NOS, LX, LY, 1, bonds)
    ref = _ref_setup(NOS, LX, LY, 1, N, (1, 1, 0), bonds)
    for t in range(1, m.comb(NOS, N) + 1):
        n = m.unrank(t, N)
        r1, r2 = m.representative(n), ref.representative(n)
        assert r1[0] == r2[0] and np.array_equal(r1[1], r2[1])


@pytest.mark.parametrize("case", ["chain", "torus"])
def test_correlations_local_mag_sq(case):
    """calcu_cf_trans_2, calcu_lm_trans_2 and calcu_Sq_a (S+, S-, Sz)."""
    if case == "chain":
        LX, LY, N, K = 10, 1, 4, (2, 0, 0)
        bonds = orc.chain_bonds(10, [(1, 1.0, 0.8), (2, 0.4, 0.3)])
    else:
        LX, LY, N, K = 4, 3, 3, (1, 1, 0)
        bonds = orc.square_bonds(4, 3, 1.0)
    NOS = LX * LY
    m = orc.Model(NOS, LX, LY, 1, bonds)
    rk = (orc.wavevector(K[0], LX), orc.wavevector(K[1], LY), 0.0)
    sec = m.build_sector(N, rk)
    ref = _ref_setup(NOS, LX, LY, 1, N, K, bonds)
    ref.build_sector()
    rng = np.random.default_rng(3)
    v = rng.uniform(-1, 1, sec.dim) + 1j * rng.uniform(-1, 1, sec.dim)
    scale = np.vdot(v, v).real
    for i in range(1, NOS + 1):
        for j in range(1, NOS + 1):
            a, b = ref.cf_pair((i, j), v), m.cf_pair(sec, v, (i, j))
            assert a[0] == b[0]
            assert abs(a[1] - b[1]) <= 1e-13 * scale and abs(a[2] - b[2]) <= 1e-13 * scale
    for n in range(1, NOS + 1):
        no, val = ref.lm_site(n, v)
        ln = m.mk_list_n(n)
        assert no == len(ln)
        if no:
            ez = m.local_mag(sec, v)
            assert abs(ez[n - 1] - val) <= 1e-13 * scale
    pos = np.array([[i % LX, i // LX, 0.0] for i in range(NOS)], dtype=float)
    q = np.array([np.pi / 3, 2 * np.pi / 3 if LY > 1 else 0.0, 0.0])
    for spsmsz in (1, 2, 3):
        Dn = ref.build_new_sector(spsmsz, q)
        lsn, lrn = ref.new_lists()
        nod_new = N - 1 if spsmsz == 1 else (N + 1 if spsmsz == 2 else N)
        new = m.build_sector(nod_new, (rk[0] + q[0], rk[1] + q[1], rk[2] + q[2]))
        assert Dn == new.dim and np.array_equal(lsn, new.list_s) and np.array_equal(lrn, new.list_r)
        fr = ref.sq(pos, v)
        fo = m.sq_a(spsmsz, sec, new, q, pos, v)
        assert np.max(np.abs(fr - fo)) <= 1e-13 * max(1.0, np.max(np.abs(fr)))
        w = rng.uniform(-1, 1, Dn) + 1j * rng.uniform(-1, 1, Dn)
        assert np.max(np.abs(ref.hv_new(w) - m.hv(new, w))) <= 1e-13 * np.max(np.abs(ref.hv_new(w)))


U1_CASES = [
    (12, 4, [(1, 1.0, 1.0)]),
    (12, 5, [(1, 1.0, 0.7), (2, 0.5, 0.5)]),
    (14, 7, [(1, 1.0, 0.8)]),
    (10, 1, [(1, 1.0, 1.0)]),
    # beyond 64 sites (the reference's lists have no such limit)
    (70, 3, [(1, 1.0, 0.9), (2, 0.4, 0.3)]),
    (130, 2, [(1, 1.0, 1.0)]),
]


@pytest.mark.skipif(not orc.RefLibU1.available(), reason="oracle/_ref U(1)-only library not built")
@pytest.mark.parametrize("L,N,fam", U1_CASES)
def test_u1_hv_against_reference(L, N, fam):
    """The oracle's U(1)-only H.v equals the reference's own
    source_only_u1/ham2vec_v3.f90 on the full C(L, N) block, and the same
    block on a 4x3 torus (bond order and the p1 < p2 swap included)."""
    for NOS, bonds in ((L, orc.chain_bonds(L, fam)), (12, orc.square_bonds(4, 3, 1.0))):
        ref = orc.RefLibU1()
        ths = ref.setup(NOS, N, bonds)
        om = orc.Model(NOS, NOS, 1, 1, bonds, ncol=N + 1)
        assert ths == om.u1_dim(N)
        rng = np.random.default_rng(3)
        v = rng.uniform(-1, 1, ths) + 1j * rng.uniform(-1, 1, ths)
        r = ref.hv(v)
        o = om.hv_u1(N, v)
        assert np.max(np.abs(o - r)) <= 1e-13 * max(1.0, np.max(np.abs(r)))


def test_reference_row_window_hv_equals_reference_hv():
    """The harness's row-window H.v (ref_hv_rows, the loop body of
    ham2vec_v3.f90:52-70 over rows i0..i1 with the reference's own routines,
    used to time the reference on a sample of a large sector) equals the
    reference's ham_to_vec_wave_vector row for row, also after the lists are
    installed from elsewhere (ref_set_lists)."""
    L, N, K = 18, 4, 5
    bonds = orc.chain_bonds(L, [(1, 1.0, 0.7)])
    ref = _ref_setup(L, L, 1, 1, N, (K, 0, 0), bonds)
    D = ref.build_sector()
    ls, lr = ref.lists()
    rng = np.random.default_rng(4)
    v = rng.uniform(-1, 1, D) + 1j * rng.uniform(-1, 1, D)
    full = ref.hv(v)
    ref.set_lists(ls, lr)
    out = np.full(D, np.nan + 0j)
    for r0, r1 in ((0, 7), (7, D // 2), (D // 2, D)):
        ref.hv_rows(v, r0, r1, out=out)
    assert np.array_equal(out, full)
