"""GPU parity of the U(1)-only path (source_only_u1): the block of every
configuration with N down spins, rows in colex rank order, no translation
symmetry.

One H.v must equal the oracle's restatement of
source_only_u1/ham2vec_v3.f90:35-65 (pinned against the reference's own
modules and against an independent bit-string block in the CPU tests) within
1e-12 relative in FP64 and 2e-6 in FP32 ((H - e_zz) v).  The lowest
eigenvalue of the U(1) block equals the lowest over the k sectors of the
translation-symmetric path (1e-10)."""
import itertools

import numpy as np
import pytest

from oracle import qs3 as orc

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _q3():
    from ed_amd import qs3 as q3
    return q3


CASES = [
    ("chain16_n8", 16, 16, 1, orc.chain_bonds(16, [(1, 1.0, 1.0)]), 8),
    ("chain18_n5_j1j2", 18, 18, 1, orc.chain_bonds(18, [(1, 1.0, 0.7), (2, 0.5, 0.3)]), 5),
    ("chain14_n3_xxz", 14, 14, 1, orc.chain_bonds(14, [(1, 0.8, 1.4)]), 3),
    ("chain20_n1", 20, 20, 1, orc.chain_bonds(20, [(1, 1.0, 1.0)]), 1),
    ("torus4x4_n6", 16, 4, 4, orc.square_bonds(4, 4, 1.0), 6),
    ("torus4x3_n5", 12, 4, 3, orc.square_bonds(4, 3, 1.0), 5),
    # beyond 64 sites: the kernel that works from the sorted down-spin sites
    ("chain70_n3_j1j2", 70, 70, 1, orc.chain_bonds(70, [(1, 1.0, 0.9), (2, 0.4, 0.3)]), 3),
    ("chain130_n2", 130, 130, 1, orc.chain_bonds(130, [(1, 1.0, 1.0)]), 2),
    ("torus9x8_n2", 72, 9, 8, orc.square_bonds(9, 8, 1.0), 2),
    ("chain65_n4", 65, 65, 1, orc.chain_bonds(65, [(1, 1.0, 0.6)]), 4),
]


def _build(case):
    q3 = _q3()
    name, NOS, LX, LY, bonds, N = case
    om = orc.Model(NOS, LX, LY, 1, bonds, ncol=max(N + 1, 2))
    gm = q3.Model(NOS, LX, LY, 1, bonds)
    gs = q3.Sector.build_u1(gm, N)
    return om, gm, gs, N


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_u1_hv_parity(case):
    print(f"[gpu] u1 hv {case[0]}", flush=True)
    q3 = _q3()
    om, gm, gs, N = _build(case)
    assert gs.dim == om.u1_dim(N)
    ls, lr = gs.lists()
    assert np.array_equal(ls, np.arange(1, gs.dim + 1)) and np.all(lr == 1.0)
    rng = np.random.default_rng(5)
    v = rng.uniform(-1, 1, gs.dim) + 1j * rng.uniform(-1, 1, gs.dim)
    ref = om.hv_u1(N, v)
    got = gs.ham_to_vec_wave_vector(v)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
    ref32 = ref - om.fm_energy() * v
    dv = torch.from_numpy(v.astype(np.complex64)).cuda()
    out = torch.empty_like(dv)
    gs.hv_dev(out.data_ptr(), dv.data_ptr(), q3.FP32)
    torch.cuda.synchronize()
    got32 = out.cpu().numpy().astype(np.complex128)
    assert np.max(np.abs(got32 - ref32)) <= 2e-6 * np.max(np.abs(ref32))
    # nnz: antiparallel bonds summed over all configurations
    name, NOS, LX, LY, bonds, _ = case
    if gs.dim <= 20000:
        nnz = 0
        for c in itertools.combinations(range(1, NOS + 1), N):
            s = set(c)
            nnz += sum(1 for (a, b, _, _) in bonds if (int(a) in s) != (int(b) in s))
        assert gs.nnz() == nnz


@pytest.mark.parametrize("prec_name", ["fp64", "fp32"])
def test_u1_ground_state_is_lowest_over_k(prec_name):
    """E0 of the U(1) block = min over k of the sector E0 (J1-J2 chain L=16,
    N=5): plain Lanczos in FP64 (1e-10) and the FP32-mixed thick restart
    (1e-9 relative)."""
    print(f"[gpu] u1 ground state {prec_name}", flush=True)
    q3 = _q3()
    L, N = 16, 5
    bonds = orc.chain_bonds(L, [(1, 1.0, 1.0), (2, 0.4, 0.4)])
    gm = q3.Model(L, L, 1, 1, bonds)
    e_k = []
    for K in range(L):
        s = q3.Sector.build(gm, N, (q3.wavevector(K, L), 0.0, 0.0))
        if s.dim:
            e, _, _ = q3.lanczos_routines(s, maxitr=2000, lnc_ene_conv=1e-14, seed=2, want_vector=False)
            e_k.append(e)
    gs = q3.Sector.build_u1(gm, N)
    if prec_name == "fp64":
        e, _, _ = q3.lanczos_routines(gs, maxitr=3000, lnc_ene_conv=1e-14, seed=2, want_vector=False)
        assert abs(e - min(e_k)) <= 1e-10 * max(1.0, abs(e)), (e, min(e_k))
    else:
        ene, psi, st = q3.my_trlanczos_routines_complex(gs, NOE=1, NOK=6, NOM=40, maxitr=400,
                                                         lnc_ene_conv=1e-11, prec=q3.FP32, seed=2)
        assert abs(ene[0] - min(e_k)) <= 1e-9 * max(1.0, abs(ene[0])), (ene[0], min(e_k), st)


def test_u1_rejects_symmetric_only_entry_points():
    q3 = _q3()
    gm = q3.Model(10, 10, 1, 1, orc.chain_bonds(10, [(1, 1.0, 1.0)]))
    gs = q3.Sector.build_u1(gm, 3)
    v = np.ones(gs.dim, dtype=np.complex128) / np.sqrt(gs.dim)
    with pytest.raises(q3.QS3XError, match="U\\(1\\)-only"):
        q3.get_cf_2_wave_vector(gs, v, [(1, 2)])


def test_u1_ground_state_beyond_64_sites():
    """L = 66 J1-J2 chain, N = 3 (C(66, 3) = 45,760 rows, the kernel that
    works from the sorted down-spin sites): E0 of the U(1) block = min over k
    of the translation-symmetric sector E0 (FP64, 1e-10)."""
    print("[gpu] u1 ground state L=66", flush=True)
    q3 = _q3()
    L, N = 66, 3
    bonds = orc.chain_bonds(L, [(1, 1.0, 1.0), (2, 0.3, 0.3)])
    gm = q3.Model(L, L, 1, 1, bonds)
    e_k = []
    for K in range(L):
        s = q3.Sector.build(gm, N, (q3.wavevector(K, L), 0.0, 0.0))
        if s.dim:
            e, _, _ = q3.lanczos_routines(s, maxitr=2000, lnc_ene_conv=1e-14, seed=2, want_vector=False)
            e_k.append(e)
    gs = q3.Sector.build_u1(gm, N)
    assert gs.dim == 45760
    e, _, _ = q3.lanczos_routines(gs, maxitr=3000, lnc_ene_conv=1e-14, seed=2, want_vector=False)
    assert gs.hv_last_launch()[0] == 4
    assert abs(e - min(e_k)) <= 1e-10 * max(1.0, abs(e)), (e, min(e_k))
