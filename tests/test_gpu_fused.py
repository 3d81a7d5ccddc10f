"""GPU parity of the fused Lanczos step on gap-lattice stencil sectors.

On nearest-neighbour chains with uniform bonds and N_down = 4, qs3x_lanczos
runs each step of the reference's recurrence (lanczos.f90:50-92: H.v,
alpha = <v|Hv>, w = Hv - alpha v - beta v', beta = |w|) as one kernel pass:
the Lanczos vector v_i = (P_{i-1} - alpha_{i-1} v_{i-1}) / beta_{i-1} is
formed from the two previous vectors as the stencil reads them, and the pass
writes v_i and P_i = H v_i - beta_{i-1} v_{i-1}; alpha_i and beta_i come from
its partial sums (beta_i^2 = |P_i|^2 - 2 alpha_i Re<v_i, P_i> +
alpha_i^2 |v_i|^2).  Checked against the unfused kernels (option
lanczos_fused 0) from the same start vector, and against dense
diagonalisation of the oracle's H on small sectors.  Tolerances: FP64 E0 to
1e-10 absolute (the engine's eigenvalue bar) and the ground state to an
overlap of 1 - 1e-8; FP32-mixed E0 to 5e-6 relative."""
import numpy as np
import pytest

from oracle import qs3 as orc

pytestmark = pytest.mark.gpu


def _q3():
    from ed_amd import qs3 as q3
    return q3


CASES = [
    # L, K, Jz, stencil lines per tile (0: default), slabs per task: N_down = 4,
    # nearest-neighbour chain
    (20, 0, 1.0, 0, 0),
    (24, 5, 0.7, 2, 3),
    (60, 0, 1.0, 4, 1),
    (160, 13, 0.8, 6, 9),
    (400, 0, 1.0, 8, 16),
]


def _run(q3, sec, prec, fused, lines=0, slabs=0, depth=0, **kw):
    # lanczos_fused 2: the fused step in both precisions (the default, 1,
    # fuses FP64 only)
    q3.set_option("lanczos_fused", 2 if fused else 0)
    q3.set_option("hv_stencil_lines", lines)
    q3.set_option("hv_stencil_slabs", slabs)
    q3.set_option("hv_stencil_depth", depth)
    try:
        E, psi, it = q3.lanczos_routines(sec, prec=prec, **kw)
        fam = sec.hv_last_launch()[0]
    finally:
        q3.set_option("lanczos_fused", 1)
        q3.set_option("hv_stencil_lines", 0)
        q3.set_option("hv_stencil_slabs", 0)
        q3.set_option("hv_stencil_depth", 0)
    return E, psi, it, fam


@pytest.mark.parametrize("L,K,jz,lines,slabs", CASES)
def test_fused_lanczos_matches_unfused_fp64(L, K, jz, lines, slabs):
    print(f"[gpu] fused lanczos FP64 L={L} K={K}", flush=True)
    q3 = _q3()
    gm = q3.Model(L, L, 1, 1, q3.chain_bonds(L, [(1, 1.0, jz)]))
    sec = q3.Sector.build(gm, 4, (q3.wavevector(K, L), 0.0, 0.0))
    kw = dict(maxitr=3000, minitr=20, itrint=5, lnc_ene_conv=1e-14, seed=7)
    E1, p1, it1, fam1 = _run(q3, sec, q3.FP64, True, lines, slabs, 3 if L % 3 == 0 else 0, **kw)
    E0, p0, it0, fam0 = _run(q3, sec, q3.FP64, False, **kw)
    assert fam1 == 6 and fam0 == 5, (fam1, fam0)
    assert abs(E1 - E0) <= 1e-10, (E1, E0)
    assert abs(abs(np.vdot(p0, p1)) - 1.0) <= 1e-8
    assert abs(np.linalg.norm(p1) - 1.0) <= 1e-12
    if sec.dim <= 3000:
        om = orc.Model(L, L, 1, 1, orc.chain_bonds(L, [(1, 1.0, jz)]))
        osec = om.build_sector(4, om.rk_from_K(K))
        w, V = np.linalg.eigh(om.full_hamiltonian(osec))
        assert abs(E1 - w[0]) <= 1e-10, (E1, w[0])
        # the residual of the returned vector under the oracle's H
        H = om.full_hamiltonian(osec)
        assert np.linalg.norm(H @ p1 - E1 * p1) <= 1e-6


@pytest.mark.parametrize("L,K,jz,lines,slabs", CASES[1:])
def test_fused_lanczos_fp32_mixed(L, K, jz, lines, slabs):
    print(f"[gpu] fused lanczos FP32 L={L} K={K}", flush=True)
    q3 = _q3()
    gm = q3.Model(L, L, 1, 1, q3.chain_bonds(L, [(1, 1.0, jz)]))
    sec = q3.Sector.build(gm, 4, (q3.wavevector(K, L), 0.0, 0.0))
    kw = dict(maxitr=3000, minitr=20, itrint=5, lnc_ene_conv=1e-12, seed=3)
    E64, p64, _, _ = _run(q3, sec, q3.FP64, True, **kw)
    E1, p1, _, fam1 = _run(q3, sec, q3.FP32, True, lines, slabs, **kw)
    E0, _, _, fam0 = _run(q3, sec, q3.FP32, False, **kw)
    assert fam1 == 6 and fam0 == 5
    for E in (E0, E1):
        assert abs(E - E64) <= 5e-6 * max(1.0, abs(E64)), (E, E64)
    assert abs(np.vdot(p64, p1)) >= 0.999


def test_fused_lanczos_c3_bethe_energy():
    """L = 1000, N_down = 4, k = 0 (D = 41,417,250): the fused FP64 and
    FP32-mixed plain Lanczos against the Bethe-ansatz E0."""
    print("[gpu] fused lanczos C3", flush=True)
    q3 = _q3()
    gm = q3.Model(1000, 1000, 1, 1, q3.chain_bonds(1000, [(1, 1.0, 1.0)]))
    sec = q3.Sector.build(gm, 4, (0.0, 0.0, 0.0))
    kw = dict(maxitr=1500, minitr=20, itrint=5, lnc_ene_conv=1e-11, seed=1, want_vector=False)
    E64, _, it64, fam = _run(q3, sec, q3.FP64, True, **kw)
    assert fam == 6
    assert abs(E64 / 242.000099091345514 - 1.0) <= 1e-9, E64
    E32, _, _, _ = _run(q3, sec, q3.FP32, True, **kw)
    assert abs(E32 / 242.000099091345514 - 1.0) <= 1e-9, E32


def test_fused_default_is_fp64_only():
    """Option lanczos_fused 1 (the default): FP64 plain Lanczos takes the
    fused step, FP32-mixed the H.v and update kernels."""
    q3 = _q3()
    gm = q3.Model(40, 40, 1, 1, q3.chain_bonds(40, [(1, 1.0, 1.0)]))
    sec = q3.Sector.build(gm, 4, (0.0, 0.0, 0.0))
    q3.lanczos_routines(sec, maxitr=30, prec=q3.FP64, want_vector=False)
    assert sec.hv_last_launch()[0] == 6
    q3.lanczos_routines(sec, maxitr=30, prec=q3.FP32, want_vector=False)
    assert sec.hv_last_launch()[0] == 5
