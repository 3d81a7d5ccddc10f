"""GPU parity of the dense ALG = 3 path: make_full_hamiltonian
(ham2vec_v3.f90:5-35) and Full_diag_routines / fulldia_routines_with_momentum
(eigen_solver.f90:51-67,263-339).

- The dense H equals the oracle's restatement of make_full_hamiltonian
  (pinned against the reference's own modules in the CPU tests) within 1e-13
  absolute, on chains and tori, real and complex phases, and on U(1)-only
  blocks (column by column against the oracle's U(1) H.v).
- The eigenvalues equal numpy's eigvalsh of the oracle's H within 1e-11; the
  eigenvectors are orthonormal and satisfy H V = V diag(E) within 1e-11 ||H||.
- Closed forms that a wrong hop sign, norm or phase would break: the one-magnon
  band of the Heisenberg chain, E = J L / 4 + J (cos 2 pi K / L - 1), in every
  k sector and as the whole U(1) block; and the reference's ALG = 3 example
  (the 6 x 6 ferromagnet of input_ex2, N_down = 3, k = 0) through the Fortran
  driver, whose ground state is the fully polarised multiplet: E0 = sum Jz / 4,
  <S^z_i S^z_j> and <S^+_i S^-_j> of the uniform superposition.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import qs3 as orc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _q3():
    from ed_amd import qs3 as q3
    return q3


SECTOR_CASES = [
    ("chain12_n4_k0_j1j2", 12, 12, 1, orc.chain_bonds(12, [(1, 1.0, 0.8), (2, 0.4, 0.3)]), 4, (0, 0)),
    ("chain12_n3_k1", 12, 12, 1, orc.chain_bonds(12, [(1, 1.0, 1.0)]), 3, (1, 0)),
    ("chain14_n5_k3_xxz", 14, 14, 1, orc.chain_bonds(14, [(1, 0.7, 1.3)]), 5, (3, 0)),
    ("torus4x4_n3_k10", 16, 4, 4, orc.square_bonds(4, 4, 1.0), 3, (1, 0)),
    ("torus4x3_n4_k12", 12, 4, 3, orc.square_bonds(4, 3, 1.0), 4, (1, 2)),
]


def _sector(case):
    q3 = _q3()
    name, NOS, LX, LY, bonds, N, K = case
    om = orc.Model(NOS, LX, LY, 1, bonds)
    gm = q3.Model(NOS, LX, LY, 1, bonds)
    rk = (orc.wavevector(K[0], LX), orc.wavevector(K[1], LY), 0.0)
    return om, om.build_sector(N, rk), q3.Sector.build(gm, N, rk)


@pytest.mark.parametrize("case", SECTOR_CASES, ids=[c[0] for c in SECTOR_CASES])
def test_full_hamiltonian_matches_oracle(case):
    print(f"[gpu] dense H {case[0]}", flush=True)
    om, osec, gs = _sector(case)
    assert gs.dim == osec.dim
    ref = om.full_hamiltonian(osec)
    got = gs.make_full_hamiltonian()
    assert np.max(np.abs(got - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref)))


@pytest.fixture
def dense_solver(request):
    q3 = _q3()
    q3.set_option("dense_solver", request.param)
    yield request.param
    q3.set_option("dense_solver", 0)


@pytest.mark.parametrize("dense_solver", [0, 1], ids=["qr", "dc"], indirect=True)
@pytest.mark.parametrize("case", SECTOR_CASES, ids=[c[0] for c in SECTOR_CASES])
def test_full_diag_eigenpairs(case, dense_solver):
    print(f"[gpu] full diag {case[0]} solver {dense_solver}", flush=True)
    om, osec, gs = _sector(case)
    H = om.full_hamiltonian(osec)
    w_ref = np.linalg.eigvalsh(0.5 * (H + H.conj().T))
    w, V = gs.full_diag()
    assert np.all(np.diff(w) >= 0)
    assert np.max(np.abs(w - w_ref)) <= 1e-11 * max(1.0, np.max(np.abs(w_ref)))
    assert np.max(np.abs(V.conj().T @ V - np.eye(gs.dim))) <= 1e-12
    hn = np.linalg.norm(H, 2)
    assert np.max(np.abs(H @ V - V * w)) <= 1e-11 * hn
    # eigenvalues only: the same values
    assert np.max(np.abs(gs.full_diag(vectors=False) - w)) <= 1e-11 * hn


def test_full_diag_u1_block_matches_oracle():
    print("[gpu] full diag u1 block", flush=True)
    q3 = _q3()
    L, N = 14, 3
    bonds = orc.chain_bonds(L, [(1, 1.0, 0.6), (2, 0.3, 0.2)])
    om = orc.Model(L, L, 1, 1, bonds, ncol=N + 1)
    gs = q3.Sector.build_u1(q3.Model(L, L, 1, 1, bonds), N)
    ref = np.array([om.hv_u1(N, e) for e in np.eye(gs.dim, dtype=np.complex128)]).T
    got = gs.make_full_hamiltonian()
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))
    w = gs.full_diag(vectors=False)
    assert np.max(np.abs(w - np.linalg.eigvalsh(ref))) <= 1e-11 * np.max(np.abs(w))


def test_one_magnon_band_closed_form():
    """N_down = 1 on the Heisenberg chain: E(K) = J L / 4 + J (cos(2 pi K / L) - 1),
    one state per k sector; the U(1) block holds the whole band."""
    print("[gpu] one-magnon band", flush=True)
    q3 = _q3()
    L, J = 10, 0.75
    bonds = orc.chain_bonds(L, [(1, J, J)])
    gm = q3.Model(L, L, 1, 1, bonds)
    band = [J * L / 4 + J * (math.cos(2 * math.pi * K / L) - 1) for K in range(L)]
    for K in range(L):
        s = q3.Sector.build(gm, 1, (q3.wavevector(K, L), 0.0, 0.0))
        w, V = s.full_diag()
        assert s.dim == 1 and abs(w[0] - band[K]) <= 1e-13 and abs(abs(V[0, 0]) - 1) <= 1e-14
    w = q3.Sector.build_u1(gm, 1).full_diag(vectors=False)
    assert np.max(np.abs(w - np.sort(band))) <= 1e-13


def _ferromagnet_6x6():
    q3 = _q3()
    bonds = orc.square_bonds(6, 6, -1.0)
    return q3.Sector.build(q3.Model(36, 6, 6, 1, bonds), 3, (0.0, 0.0, 0.0))


def test_full_diag_degenerate_ferromagnet():
    """The 6 x 6 Heisenberg ferromagnet (N_down = 3, k = 0) has exactly
    degenerate multiplets; the default solver returns an orthonormal,
    finite eigenbasis with E0 = sum Jz / 4 = -18 once."""
    print("[gpu] full diag degenerate ferromagnet", flush=True)
    s = _ferromagnet_6x6()
    w, V = s.full_diag()
    H = s.make_full_hamiltonian()
    assert np.all(np.isfinite(V)) and abs(w[0] + 18) <= 1e-12 and w[1] + 18 > 1e-3
    assert np.max(np.abs(V.conj().T @ V - np.eye(s.dim))) <= 1e-12
    assert np.max(np.abs(H @ V - V * w)) <= 1e-11 * np.linalg.norm(H, 2)


@pytest.mark.parametrize("dense_solver", [1], ids=["dc"], indirect=True)
def test_full_diag_divide_and_conquer_is_checked(dense_solver):
    """With divide and conquer the result on the same spectrum is either a
    correct eigenbasis or an error that names non-finite output: never NaN
    returned as a result."""
    print("[gpu] full diag zheevd checked", flush=True)
    q3 = _q3()
    s = _ferromagnet_6x6()
    try:
        w, V = s.full_diag()
    except q3.QS3XError as e:
        assert "non-finite" in str(e)
        return
    assert np.all(np.isfinite(V)) and abs(w[0] + 18) <= 1e-12
    assert np.max(np.abs(V.conj().T @ V - np.eye(s.dim))) <= 1e-12


def test_full_diag_refuses_the_all_k_sector():
    q3 = _q3()
    gm = q3.Model(12, 12, 1, 1, orc.chain_bonds(12, [(1, 1.0, 1.0)]))
    with pytest.raises(q3.QS3XError, match="all-k"):
        q3.Sector.build_all_k(gm, 3).full_diag()


def _write_ex2_deck(d):
    """input_ex2's deck (6 x 6 square lattice, J = -1, N_down = 3, k = 0,
    ALG = 3, lm, cf and S(q, w) on), with the bond and pair lists that
    input_ex1 holds, written here."""
    from ed_amd import inputs as inp
    os.makedirs(os.path.join(d, "out"), exist_ok=True)
    bonds = orc.square_bonds(6, 6, -1.0)
    inp.write_bond_list(os.path.join(d, "bonds.dat"), bonds)
    with open(os.path.join(d, "lm.dat"), "w") as f:
        f.write("".join(f"{i}\n" for i in range(1, 37)))
    with open(os.path.join(d, "cf.dat"), "w") as f:
        f.write("".join(f"{i:5d}{j:5d}\n" for i in range(1, 37) for j in range(1, 37)))
    inp.write_positions(os.path.join(d, "pos.dat"), 6, 6, 1)
    inp.write_input_dat(
        os.path.join(d, "input.dat"),
        input_parameters=dict(NOS=36, NOD=3, LX=6, LY=6, LZ=1, KX=0, KY=0, KZ=0, NOxxz=len(bonds), ALG=3,
                              cal_lm=1, cal_cf=1, cal_dsf=1, wr_wf=0, re_wf=0, FILExxz="bonds.dat",
                              FILEwf="out/", OUTDIR="out/"),
        input_static=dict(NOV=2, NOLM=36, NOCF=1296, FILElm="lm.dat", FILECF="cf.dat"),
        input_dynamic=dict(spsmsz=3, itr_dsf=200, QX=0.0, QY=0.0, QZ=0.0, rfield=0.495, FILEpos="pos.dat"),
        input_lancz=dict(lnc_ene_conv=1e-14, minitr=20, maxitr=10000, itrint=5),
        input_TRLan=dict(NOE=10, NOK=15, NOM=30, maxitr=10000, lnc_ene_conv=1e-14, i_vec_min=1, i_vec_max=1))
    return bonds


def _cols(path):
    return np.array([[float(x) for x in ln.split()] for ln in open(path) if ln.strip()])


def test_driver_alg3_reference_example(tmp_path):
    print("[gpu] fortran driver ALG=3 (6x6 ferromagnet, N_down=3)", flush=True)
    d = str(tmp_path)
    bonds = _write_ex2_deck(d)
    exe = os.path.join(ROOT, "ed_amd", "fortran", "qs3x_main.exe")
    with open(os.path.join(d, "input.dat")) as fin:
        r = subprocess.run([exe], stdin=fin, cwd=d, capture_output=True, text=True, timeout=110)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    L, N = 36, 3
    om = orc.Model(L, 6, 6, 1, bonds)
    sec = om.build_sector(N, (0.0, 0.0, 0.0))
    w_ref = np.linalg.eigvalsh(om.full_hamiltonian(sec))
    w = _cols(os.path.join(d, "out", "full_eigenvalues.dat"))[:, 0]
    assert len(w) == sec.dim and np.max(np.abs(w - w_ref)) <= 1e-11
    e_fm = sum(b[3] for b in bonds) / 4
    assert abs(w[0] - e_fm) <= 1e-12 and w[1] - w[0] > 1e-3
    # every eigenvector is translation invariant: <S^z_j> = (L/2 - N)/L, all nvec = dim columns
    lm = _cols(os.path.join(d, "out", "local_mag.dat"))
    assert lm.shape == (L, 1 + sec.dim)
    assert np.max(np.abs(lm[:, 1:] - (L / 2 - N) / L)) <= 1e-12
    # the ground state is the uniform superposition of all C(L, N) configurations
    szsz = _cols(os.path.join(d, "out", "SzSz-corr.dat"))
    spsm = _cols(os.path.join(d, "out", "SpSm-corr.dat"))
    i, j = szsz[:, 0].astype(int), szsz[:, 1].astype(int)
    same = i == j
    zz = np.where(same, 0.25, 0.25 - N / L + N * (N - 1) / (L * (L - 1)))
    pm = np.where(same, 1 - N / L, math.comb(L - 2, N - 1) / math.comb(L, N))
    assert np.max(np.abs(szsz[:, 2] - zz)) <= 1e-12
    assert np.max(np.abs(spsm[:, 2] - pm)) <= 1e-12 and np.max(np.abs(spsm[:, 3])) <= 1e-12
    # the ground-state vector file: one complex amplitude per row
    amps = open(os.path.join(d, "out", "full_eigenvectors.dat")).read().split("\n")
    assert len([a for a in amps if a.strip()]) == sec.dim
    ab = [f for f in os.listdir(os.path.join(d, "out")) if f.startswith("alpha_beta_values_")]
    assert ab, "S(q, w) continued-fraction coefficients were not written"
