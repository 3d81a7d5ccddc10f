"""GPU parity: solvers, Ritz rotation, observables and the DSF recurrence.

Tolerances (from the engine's contract): eigenvalues 1e-10 (absolute, energies
are O(1)-O(L)); correlation functions and continued-fraction coefficients 1e-8;
FP32-mixed results are stated against the FP64 ones in the assertion."""
import ctypes
import math

import numpy as np
import pytest

from oracle import qs3 as orc

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _q3():
    from ed_amd import qs3 as q3
    return q3


def _setup(L, N, fam, K=0, LY=1):
    q3 = _q3()
    if LY == 1:
        bonds = orc.chain_bonds(L, fam)
        om = orc.Model(L, L, 1, 1, bonds)
        gm = q3.Model(L, L, 1, 1, bonds)
        rk = om.rk_from_K(K)
    else:
        LX = L // LY
        bonds = orc.square_bonds(LX, LY, 1.0)
        om = orc.Model(L, LX, LY, 1, bonds)
        gm = q3.Model(L, LX, LY, 1, bonds)
        rk = (orc.wavevector(K[0], LX), orc.wavevector(K[1], LY), 0.0)
    return om, om.build_sector(N, rk), gm, q3.Sector.build(gm, N, rk), rk


def test_start_vector_generator_matches_formula():
    print("[gpu] start vector", flush=True)
    q3 = _q3()
    om, osec, gm, gsec, rk = _setup(12, 4, [(1, 1.0, 1.0)])
    # one Lanczos step with maxitr=1 returns alpha = <v|H|v> of the generated vector
    v = q3.start_vector(gsec.dim, 77)
    v = v / np.linalg.norm(v)
    H = om.full_hamiltonian(osec)
    E, psi, it = q3.lanczos_routines(gsec, maxitr=1, minitr=1, itrint=1, seed=77, want_vector=True)
    assert it == 1 and E == pytest.approx(np.vdot(v, H @ v).real, abs=1e-12)
    # after one step the Ritz vector is the normalised start vector itself:
    # the device generator equals the host formula element by element
    assert np.max(np.abs(psi - v)) <= 1e-14, np.max(np.abs(psi - v))
    # and a host start vector equal to it gives the same energy bit for bit
    E2, _, _ = q3.lanczos_routines(gsec, maxitr=1, minitr=1, itrint=1, start=q3.start_vector(gsec.dim, 77),
                                   want_vector=False)
    assert E2 == E


@pytest.mark.parametrize("prec_name", ["fp64", "fp32"])
def test_plain_lanczos(prec_name):
    print(f"[gpu] plain lanczos {prec_name}", flush=True)
    q3 = _q3()
    prec = q3.FP64 if prec_name == "fp64" else q3.FP32
    om, osec, gm, gsec, rk = _setup(16, 6, [(1, 1.0, 1.0), (2, 0.4, 0.4)], K=2)
    w, V = np.linalg.eigh(om.full_hamiltonian(osec))
    E, psi, it = q3.lanczos_routines(gsec, maxitr=2000, minitr=20, itrint=5, lnc_ene_conv=1e-14, prec=prec, seed=3)
    tol = 1e-10 if prec == q3.FP64 else 5e-6 * max(1.0, abs(w[0]))
    assert E == pytest.approx(w[0], abs=tol)
    ov = abs(np.vdot(V[:, 0], psi))
    assert ov > (1 - 1e-8 if prec == q3.FP64 else 0.999)


def test_lanczos_bethe_c1_size():
    """C1 (L=36, N=6, k=0, D=54,132): E0 against the Bethe-ansatz value."""
    print("[gpu] lanczos C1", flush=True)
    q3 = _q3()
    gm = q3.Model(36, 36, 1, 1, q3.chain_bonds(36, [(1, 1.0, 1.0)]))
    s = q3.Sector.build(gm, 6, (0.0, 0.0, 0.0))
    assert s.dim == 54132
    E, _, it = q3.lanczos_routines(s, maxitr=3000, lnc_ene_conv=1e-14, seed=1, want_vector=False)
    assert E == pytest.approx(-2.687328296769200, abs=1e-12)


@pytest.mark.parametrize("prec_name", ["fp64", "fp32"])
def test_thick_restart(prec_name):
    print(f"[gpu] thick restart {prec_name}", flush=True)
    q3 = _q3()
    prec = q3.FP64 if prec_name == "fp64" else q3.FP32
    om, osec, gm, gsec, rk = _setup(18, 6, [(1, 1.0, 1.0), (2, 0.5, 0.5)], K=0)
    w, V = np.linalg.eigh(om.full_hamiltonian(osec))
    ene, psi, st = q3.my_trlanczos_routines_complex(gsec, NOE=4, NOK=8, NOM=24, maxitr=500, lnc_ene_conv=1e-13,
                                                     prec=prec, seed=5)
    # eigenvalues within 1e-10 of dense diagonalisation (FP32 path refines in FP64)
    assert np.allclose(ene, w[:4], atol=1e-10), (ene, w[:4])
    H = om.full_hamiltonian(osec)
    for j in range(4):
        r = np.linalg.norm(H @ psi[:, j] - ene[j] * psi[:, j])
        assert r < 1e-5, r


@pytest.mark.parametrize("nom,reorth,prec_name", [(40, 1, "fp32"), (150, 1, "fp32"), (150, 2, "fp32"),
                                                   (150, 2, "fp64")])
def test_thick_restart_checked(nom, reorth, prec_name):
    """Convergence checks inside the cycles (tr_check_every), three-term first
    cycle (tr_reorth 1; 2: three-term steps in restart cycles too), long
    cycles (NOM 150 uses the structured projected eigensolver); in FP32 then
    the FP64 Rayleigh-Ritz step and refinement."""
    print(f"[gpu] thick restart checked NOM={nom} reorth={reorth} {prec_name}", flush=True)
    q3 = _q3()
    prec = q3.FP64 if prec_name == "fp64" else q3.FP32
    om, osec, gm, gsec, rk = _setup(20, 5, [(1, 1.0, 1.0)], K=0)
    w = np.linalg.eigvalsh(om.full_hamiltonian(osec))
    q3.set_option("tr_check_every", 5)
    q3.set_option("tr_reorth", reorth)
    q3.set_option("tr_cgs_passes", 1)
    try:
        ene, psi, st = q3.my_trlanczos_routines_complex(gsec, NOE=1, NOK=4, NOM=nom, maxitr=200,
                                                         lnc_ene_conv=1e-12, prec=prec, seed=3)
    finally:
        q3.set_option("tr_check_every", 0)
        q3.set_option("tr_reorth", 0)
        q3.set_option("tr_cgs_passes", 2)
    assert abs(ene[0] - w[0]) <= 1e-9 * abs(w[0]), (ene[0], w[0], st)
    H = om.full_hamiltonian(osec)
    r = np.linalg.norm(H @ psi[:, 0] - ene[0] * psi[:, 0])
    assert r < 1e-4, r


@pytest.mark.parametrize("rot_vec", [1, 0])
def test_ritz_rotation_mfma(rot_vec):
    """The FP64 MFMA rotation equals numpy's V[:, :nom] @ C for both storages,
    with 16-byte basis loads (aligned slots: every FP64 case, even dims in
    FP32) and with element loads."""
    print(f"[gpu] ritz rotation rot_vec={rot_vec}", flush=True)
    q3 = _q3()
    q3.set_option("rot_vec", rot_vec)
    L = q3.lib()
    L.qs3x_test_ritz_rotate.restype = ctypes.c_int
    L.qs3x_test_ritz_rotate.argtypes = [ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32,
                                        ctypes.c_int32, ctypes.c_void_p]
    rng = np.random.default_rng(9)
    for dt, tol in ((np.complex128, 1e-13), (np.complex64, 2e-6)):
        # the last two stage 94-113 KB of coefficients in LDS (beyond the default 64 KB)
        for dim, nom, nok in ((1000, 12, 5), (777, 30, 17), (5, 8, 4), (4099, 41, 33), (4098, 64, 64), (130, 96, 48),
                              (3000, 200, 64), (2001, 300, 40)):
            V = (rng.normal(size=(nom + 1, dim)) + 1j * rng.normal(size=(nom + 1, dim))).astype(dt)
            C = rng.normal(size=(nom, nok))
            ref = (C.T @ V[:nom].astype(np.complex128))
            d = torch.from_numpy(V.copy()).cuda()
            Cc = np.ascontiguousarray(C)
            rc = L.qs3x_test_ritz_rotate(0 if dt == np.complex128 else 1, ctypes.c_void_p(d.data_ptr()), dim, nom,
                                         nok, ctypes.c_void_p(Cc.ctypes.data))
            assert rc == 0
            got = d.cpu().numpy()
            assert np.max(np.abs(got[:nok] - ref)) <= tol * np.max(np.abs(ref))
            assert np.array_equal(got[nok:], V[nok:])  # untouched columns
    q3.set_option("rot_vec", 1)


@pytest.mark.parametrize("geom", ["chain", "torus"])
def test_correlations_and_local_mag(geom):
    print(f"[gpu] correlations {geom}", flush=True)
    q3 = _q3()
    if geom == "chain":
        om, osec, gm, gsec, rk = _setup(12, 5, [(1, 1.0, 0.8), (2, 0.4, 0.3)], K=3)
    else:
        om, osec, gm, gsec, rk = _setup(12, 3, None, K=(1, 1), LY=3)
    rng = np.random.default_rng(4)
    v = rng.uniform(-1, 1, osec.dim) + 1j * rng.uniform(-1, 1, osec.dim)
    NOS = om.NOS
    pairs = [(a, b) for a in range(1, NOS + 1) for b in range(1, NOS + 1)]
    sz_o, sp_o = om.correlations(osec, v, pairs)
    sz_g, sp_g = q3.get_cf_2_wave_vector(gsec, v, pairs)
    scale = np.vdot(v, v).real
    assert np.max(np.abs(sz_o - sz_g)) <= 1e-12 * scale
    assert np.max(np.abs(sp_o - sp_g)) <= 1e-12 * scale
    lm_o = om.local_mag(osec, v)
    lm_g = q3.get_lm_2_wave_vector(gsec, v)
    assert np.max(np.abs(lm_o - lm_g)) <= 1e-12 * scale


@pytest.mark.parametrize("geom,spsmsz", [("chain", 1), ("chain", 2), ("chain", 3), ("torus", 1), ("torus", 2),
                                         ("torus", 3)])
def test_apply_sq(geom, spsmsz):
    print(f"[gpu] S_q {geom} {spsmsz}", flush=True)
    q3 = _q3()
    if geom == "chain":
        om, osec, gm, gsec, rk = _setup(12, 4, [(1, 1.0, 1.0)], K=2)
        q = np.array([2 * np.pi * 5 / 12, 0.0, 0.0])
        LX, LY = 12, 1
    else:
        om, osec, gm, gsec, rk = _setup(12, 3, None, K=(1, 2), LY=3)
        q = np.array([np.pi / 2, 2 * np.pi / 3, 0.0])
        LX, LY = 4, 3
    NOS = om.NOS
    pos = np.array([[i % LX, i // LX, 0.0] for i in range(NOS)], dtype=float)
    dn = {1: -1, 2: 1, 3: 0}[spsmsz]
    rkn = (rk[0] + q[0], rk[1] + q[1], rk[2] + q[2])
    onew = om.build_sector(osec.nod + dn, rkn)
    gnew = q3.Sector.build(gm, osec.nod + dn, rkn)
    rng = np.random.default_rng(6)
    v = rng.uniform(-1, 1, osec.dim) + 1j * rng.uniform(-1, 1, osec.dim)
    fo = om.sq_a(spsmsz, osec, onew, q, pos, v)
    fg = q3.calcu_Sq_a(spsmsz, gsec, gnew, q, pos, v)
    assert np.max(np.abs(fo - fg)) <= 1e-12 * max(1.0, np.max(np.abs(fo)))


@pytest.mark.parametrize("prec_name", ["fp64", "fp32"])
def test_dsf_recurrence(prec_name):
    print(f"[gpu] dsf {prec_name}", flush=True)
    q3 = _q3()
    prec = q3.FP64 if prec_name == "fp64" else q3.FP32
    om, osec, gm, gsec, rk = _setup(14, 5, [(1, 1.0, 1.0)], K=0)
    rng = np.random.default_rng(8)
    f = rng.uniform(-1, 1, osec.dim) + 1j * rng.uniform(-1, 1, osec.dim)
    H = om.full_hamiltonian(osec)
    sztot_new = 14 / 2.0 - 5
    a_o, b_o, n_o = orc.dsf_recurrence(lambda x: H @ x - 0.3 * sztot_new * x, f, 40)
    a_g, b_g, n_g = q3.dsf_recurrence(gsec, f, 40, rfield=0.3, prec=prec)
    assert n_g == pytest.approx(n_o, rel=1e-12)
    tol = 1e-8 if prec == q3.FP64 else 1e-4
    # early coefficients: loss of orthogonality makes late ones chaotic in any precision
    assert np.allclose(a_g[:20], a_o[:20], atol=tol * max(1, np.max(np.abs(a_o[:20]))))
    assert np.allclose(b_g[:20], b_o[:20], atol=tol * max(1, np.max(np.abs(b_o[:20]))))
    om_ = np.linspace(0, 4, 201)
    s_o = orc.continued_fraction(a_o, b_o, n_o, 0.0, om_, 0.05)
    s_g = q3.dsf_analysis(a_o, b_o, n_o, 0.0, om_, 0.05)
    assert np.allclose(s_o, s_g, rtol=1e-12, atol=1e-14)


def test_fp32_mixed_solvers_large_chain():
    """FP32-mixed solvers on an 83,360-row sector (chain L=128, N_down=4,
    k=0: the two-rows-per-lane H.v and the 16-byte Lanczos update) against
    plain FP64 Lanczos: plain FP32-mixed Lanczos within 1e-8 relative, and the
    FP32-mixed thick restart with its FP64 Rayleigh-Ritz step and refinement
    within 1e-9."""
    print("[gpu] fp32-mixed solvers L=128 N=4", flush=True)
    q3 = _q3()
    gm = q3.Model(128, 128, 1, 1, q3.chain_bonds(128, [(1, 1.0, 1.0)]))
    s = q3.Sector.build(gm, 4, (0.0, 0.0, 0.0))
    assert s.dim == 83360
    E64, _, _ = q3.lanczos_routines(s, maxitr=3000, minitr=20, itrint=5, lnc_ene_conv=1e-14, prec=q3.FP64, seed=1,
                                    want_vector=False)
    E32, _, _ = q3.lanczos_routines(s, maxitr=3000, minitr=20, itrint=5, lnc_ene_conv=1e-12, prec=q3.FP32, seed=1,
                                    want_vector=False)
    assert abs(E32 - E64) <= 1e-8 * abs(E64), (E32, E64)
    q3.set_option("tr_check_every", 5)
    q3.set_option("tr_reorth", 2)
    q3.set_option("tr_cgs_passes", 1)
    try:
        ene, _, st = q3.my_trlanczos_routines_complex(s, NOE=1, NOK=4, NOM=120, maxitr=400, lnc_ene_conv=1e-11,
                                                       prec=q3.FP32, seed=1, want_vectors=False)
    finally:
        q3.set_option("tr_check_every", 0)
        q3.set_option("tr_reorth", 0)
        q3.set_option("tr_cgs_passes", 2)
    assert abs(ene[0] - E64) <= 1e-9 * abs(E64), (ene[0], E64, st)


@pytest.mark.parametrize("prec_name", ["fp64", "fp32"])
def test_thick_restart_many_restarts_small_basis(prec_name):
    """Thick restart with the three-term first cycle (tr_reorth 2), one
    Gram-Schmidt pass requested and checks inside cycles, on a basis small
    enough to force many restarts (chain L = 200, N_down = 4, NOM 24): the
    restart cycles must keep the new vectors orthogonal to the locked Ritz
    vectors (eigen_solver.f90:233-239 reorthogonalises fully), so E0 equals
    plain FP64 Lanczos (lanczos.f90:50-152) to 1e-10 (FP64) / 1e-8 (FP32-mixed)."""
    print(f"[gpu] thick restart many restarts {prec_name}", flush=True)
    q3 = _q3()
    prec = q3.FP64 if prec_name == "fp64" else q3.FP32
    gm = q3.Model(200, 200, 1, 1, q3.chain_bonds(200, [(1, 1.0, 1.0)]))
    s = q3.Sector.build(gm, 4, (0.0, 0.0, 0.0))
    E64, _, _ = q3.lanczos_routines(s, maxitr=4000, lnc_ene_conv=1e-14, prec=q3.FP64, seed=1, want_vector=False)
    q3.set_option("tr_check_every", 5)
    q3.set_option("tr_reorth", 2)
    q3.set_option("tr_cgs_passes", 1)
    try:
        ene, _, st = q3.my_trlanczos_routines_complex(s, NOE=1, NOK=4, NOM=24, maxitr=500, lnc_ene_conv=1e-12,
                                                       prec=prec, seed=1, want_vectors=False)
    finally:
        q3.set_option("tr_check_every", 0)
        q3.set_option("tr_reorth", 0)
        q3.set_option("tr_cgs_passes", 2)
    assert st["restarts"] >= 5
    tol = 1e-10 if prec == q3.FP64 else 1e-8
    assert abs(ene[0] - E64) <= tol * abs(E64), (ene[0], E64, st)


@pytest.mark.parametrize("L,N,K,fam", [(16, 6, 2, [(1, 1.0, 1.0), (2, 0.4, 0.4)]), (24, 4, 0, [(1, 1.0, 0.8)]),
                                       (30, 5, 3, [(1, 1.0, 1.0)])])
def test_fp32_mixed_fp64_lanczos_certificate(L, N, K, fam):
    """Option fp64_lanczos: the FP32-mixed thick restart ends with plain FP64
    Lanczos (lanczos.f90:50-92) from its Rayleigh-Ritz vector until the
    residual bound r^2 / gap certifies E0 (eigen_solver.f90:150-162 is the
    reference's stop test).  The value and the returned vector are checked
    against dense diagonalisation of the oracle's H: E0 to 1e-10, the vector's
    residual |H psi - E psi| <= 1e-5, and the certificate is the one reported."""
    print(f"[gpu] fp64 lanczos certificate L={L} N={N} K={K}", flush=True)
    q3 = _q3()
    om, osec, gm, gsec, rk = _setup(L, N, fam, K=K)
    H = om.full_hamiltonian(osec)
    w = np.linalg.eigvalsh(H)
    q3.set_option("fp64_lanczos", 300)
    q3.set_option("fp64_refine", 0)
    try:
        ene, psi, st = q3.my_trlanczos_routines_complex(gsec, NOE=1, NOK=4, NOM=40, maxitr=400, lnc_ene_conv=1e-12,
                                                         prec=q3.FP32, seed=3)
        cert = q3.solver_certificate()
    finally:
        q3.set_option("fp64_lanczos", 0)
        q3.set_option("fp64_refine", 4)
    assert abs(ene[0] - w[0]) <= 1e-10 * max(1.0, abs(w[0])), (ene[0], w[0])
    v = psi[:, 0] if psi.ndim == 2 else psi
    assert abs(np.linalg.norm(v) - 1.0) <= 1e-10
    assert np.linalg.norm(H @ v - ene[0] * v) <= 1e-5
    assert cert["accepted"] == 1.0 and cert["bound"] <= 1e-12 * abs(ene[0]) + 1e-300
