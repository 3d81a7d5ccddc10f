"""The J1-J2 chain configuration: periodic XXZ chain, L=48, J1=1 on (i, i+1),
J2=0.5 on (i, i+2) (96 bonds), N_down=6, k=0 (D = 255,704).  Thick restart
for the 4 lowest eigenvalues, then <S^z_1 S^z_j> and <S^+_1 S^-_j> for every
j, in FP64 and in FP32-mixed (FP32 vectors, FP64 alpha/beta, FP64
Rayleigh-Ritz and the final FP64 block-Krylov refinement).

What pins the results:
- the sector lists equal the oracle's (bit-exact) and one H.v equals the
  oracle's within 1e-12 relative;
- the FP64 thick-restart E0 equals the FP64 plain-Lanczos E0 within 1e-10
  (two solvers over the pinned operator), and every returned Ritz pair has a
  residual |H psi - E psi| below 1e-8 with the oracle's own H.v (a residual r
  bounds the distance to the spectrum by r);
- the FP64 correlations equal the oracle's on the same vector within 1e-12.
FP32-mixed against FP64: eigenvalues within 1e-10 relative, correlations
within 1e-8; the measured errors are printed."""
import numpy as np
import pytest

from oracle import qs3 as orc

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

L, N = 48, 6
FAM = [(1, 1.0, 1.0), (2, 0.5, 0.5)]


@pytest.fixture(scope="module")
def c2():
    from ed_amd import qs3 as q3
    bonds = orc.chain_bonds(L, FAM)
    om = orc.Model(L, L, 1, 1, bonds)
    osec = om.build_sector(N, (0.0, 0.0, 0.0))
    gm = q3.Model(L, L, 1, 1, bonds)
    gsec = q3.Sector.build(gm, N, (0.0, 0.0, 0.0))
    return q3, om, osec, gsec


def test_c2_sector_and_hv(c2):
    print("[gpu] J1-J2 L=48 N=6: sector and H.v", flush=True)
    q3, om, osec, gsec = c2
    assert gsec.dim == osec.dim == 255704
    ls, lr = gsec.lists()
    assert np.array_equal(ls, osec.list_s) and np.array_equal(lr, osec.list_r)
    rng = np.random.default_rng(7)
    v = rng.uniform(-1, 1, osec.dim) + 1j * rng.uniform(-1, 1, osec.dim)
    ref = om.hv(osec, v)
    got = gsec.ham_to_vec_wave_vector(v)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_c2_thick_restart_and_correlations(c2):
    print("[gpu] J1-J2 L=48 N=6: thick restart FP64 / FP32-mixed, correlations", flush=True)
    q3, om, osec, gsec = c2
    e64, psi64, st64 = q3.my_trlanczos_routines_complex(gsec, NOE=4, NOK=8, NOM=48, maxitr=500,
                                                         lnc_ene_conv=1e-12, prec=q3.FP64, seed=3)
    el, _, _ = q3.lanczos_routines(gsec, maxitr=3000, lnc_ene_conv=1e-14, prec=q3.FP64, seed=3, want_vector=False)
    assert abs(e64[0] - el) <= 1e-10 * max(1.0, abs(el)), (e64[0], el)
    assert np.all(np.diff(e64) >= -1e-12)
    for j in range(4):
        r = np.linalg.norm(om.hv(osec, psi64[:, j]) - e64[j] * psi64[:, j])
        assert r < 1e-8, (j, r)
    e32, psi32, st32 = q3.my_trlanczos_routines_complex(gsec, NOE=4, NOK=8, NOM=48, maxitr=500,
                                                         lnc_ene_conv=1e-10, prec=q3.FP32, seed=3)
    rel = np.abs(e32 - e64) / np.maximum(1.0, np.abs(e64))
    print(f"[gpu]   E64 {e64}  FP32-mixed rel. err {rel}  stats {st32}", flush=True)
    assert np.all(rel <= 1e-10), rel
    # the certificate of the lowest pair (Kato-Temple: r^2 / gap, gap from the
    # FP32 projection's next level outside the pair's cluster) bounds its
    # actual error; pairs whose bound misses conv continue in FP64
    cert = q3.solver_certificate()
    assert cert["gap"] > 1e-3 and cert["accepted"] in (0.0, 1.0, 2.0), cert
    assert abs(e32[0] - e64[0]) <= cert["bound"] + 1e-15, (abs(e32[0] - e64[0]), cert)
    assert st32["rr_accepted"] == cert["accepted"]
    pairs = [(1, j) for j in range(1, L + 1)]
    sz64, sp64 = q3.get_cf_2_wave_vector(gsec, psi64[:, 0], pairs)
    sz_o, sp_o = om.correlations(osec, psi64[:, 0], pairs)
    assert np.max(np.abs(sz64 - sz_o)) <= 1e-12 and np.max(np.abs(sp64 - sp_o)) <= 1e-12
    # the ground state is non-degenerate here, so its correlations are
    # independent of the eigenvector's phase and of the solver
    assert e64[1] - e64[0] > 1e-3
    sz32, sp32 = q3.get_cf_2_wave_vector(gsec, psi32[:, 0], pairs)
    dz, dp = np.max(np.abs(sz32 - sz64)), np.max(np.abs(sp32 - sp64))
    print(f"[gpu]   correlations FP32-mixed vs FP64: SzSz {dz:.2e}, S+S- {dp:.2e}", flush=True)
    assert dz <= 1e-8 and dp <= 1e-8, (dz, dp)
