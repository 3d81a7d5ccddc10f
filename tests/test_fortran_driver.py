"""The Fortran boundary: ISO_C_BINDING shims with QS3's module interfaces.

CPU: the reference's own driver (source/main.f90) compiles and links against
the shim modules unchanged (skipped where the reference sources are absent).
GPU: the engine's driver runs input decks in the reference's formats and its
output files agree with the oracle; the reference's own driver, built against
the shim modules by oracle/build_ref.sh, runs the same decks and writes the
same files and values."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import qs3 as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "ed_amd", "fortran")
FC = "/opt/rocm/lib/llvm/bin/flang"
REF_MAIN = os.path.join(os.environ.get("QS3_REFERENCE_ROOT", "/root/reference"), "source", "main.f90")


@pytest.mark.skipif(not (os.path.exists(REF_MAIN) and os.path.exists(FC)), reason="reference or flang absent")
def test_reference_driver_links_against_shims(tmp_path):
    build = os.path.join(FDIR, "_build")
    assert os.path.exists(os.path.join(build, "qs3x_modules.o")), "run ed_amd/fortran/build_fortran.sh"
    subprocess.run([FC, "-O2", "-fdefault-integer-8", "-I", build, "-c", REF_MAIN, "-o", str(tmp_path / "m.o")],
                   check=True)
    subprocess.run([FC, "-o", str(tmp_path / "x"), str(tmp_path / "m.o"), os.path.join(build, "qs3x_modules.o"),
                    os.path.join(build, "qs3x_iso_c.o"), "-L" + os.path.join(ROOT, "ed_amd"), "-lqs3x"], check=True)
    assert os.path.exists(tmp_path / "x")


def _write_deck(d, NOS, LX, LY, nod, K, bonds, alg, cal_dsf=0, spsmsz=2, itr_dsf=30, q=(0.0, 0.0, 0.0),
                rfield=0.0, nov=2, trlan=(3, 6, 16), wr_wf=0, re_wf=0, outdir="out/", filewf="out/"):
    os.makedirs(os.path.join(d, outdir), exist_ok=True)
    os.makedirs(os.path.join(d, filewf), exist_ok=True)
    with open(os.path.join(d, "bonds.dat"), "w") as f:
        for (a, b, jxy, jz) in bonds:
            f.write(f"{int(a):5d}{int(b):5d} {jxy: .15E} {jz: .15E}\n")
    with open(os.path.join(d, "lm.dat"), "w") as f:
        for i in range(1, NOS + 1):
            f.write(f"{i}\n")
    with open(os.path.join(d, "cf.dat"), "w") as f:
        for i in range(1, NOS + 1):
            for j in range(1, NOS + 1):
                f.write(f"{i:5d}{j:5d}\n")
    with open(os.path.join(d, "pos.dat"), "w") as f:
        for s in range(NOS):
            f.write(f"{float(s % LX): .15E} {float(s // LX): .15E} {0.0: .15E}\n")
    noe, nok, nom = trlan
    deck = f"""&input_parameters
NOS = {NOS}, NOD = {nod}, LX = {LX}, LY = {LY}, LZ = 1, KX = {K[0]}, KY = {K[1]}, KZ = 0,
NOxxz = {len(bonds)}, ALG = {alg}, cal_lm = 1, cal_cf = 1, cal_dsf = {cal_dsf}, wr_wf = {wr_wf}, re_wf = {re_wf},
FILExxz = "bonds.dat", FILEwf = "{filewf}", OUTDIR = "{outdir}",
&end
&input_static
NOV = {nov}, NOLM = {NOS}, NOCF = {NOS * NOS}, FILElm = "lm.dat", FILECF = "cf.dat",
&end
&input_dynamic
spsmsz = {spsmsz}, itr_dsf = {itr_dsf}, QX = {q[0]:.17f}d0, QY = {q[1]:.17f}d0, QZ = 0.0d0, rfield = {rfield}d0,
FILEpos = "pos.dat",
&end
&input_lancz
lnc_ene_conv = 1.0d-14, minitr = 20, maxitr = 2000, itrint = 5,
&end
&input_TRLan
NOE = {noe}, NOK = {nok}, NOM = {nom}, maxitr = 500, lnc_ene_conv = 1.0d-13, i_vec_min = 1, i_vec_max = 1,
&end
"""
    with open(os.path.join(d, "input.dat"), "w") as f:
        f.write(deck)


REF_DRIVER = os.path.join(ROOT, "oracle", "_ref", "qs3_main_on_qs3x.exe")


def _run(d, expect_ok=True, exe=None):
    exe = exe or os.path.join(FDIR, "qs3x_main.exe")
    assert os.path.exists(exe), "run ed_amd/fortran/build_fortran.sh"
    with open(os.path.join(d, "input.dat")) as fin:
        r = subprocess.run([exe], stdin=fin, cwd=d, capture_output=True, text=True, timeout=100)
    if not expect_ok:
        return r
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return r.stdout


def _read_cols(path):
    rows = []
    with open(path) as f:
        for line in f:
            if line.strip():
                rows.append([float(x) for x in line.split()])
    return np.array(rows)


@pytest.mark.gpu
def test_driver_thick_restart_square_lattice(tmp_path):
    """ALG=2 on a 4x4 torus (N=3, k=(1,0)): energies, <S^z>, correlations."""
    print("[gpu] fortran driver ALG=2", flush=True)
    d = str(tmp_path)
    LX, LY, nod, K = 4, 4, 3, (1, 0)
    bonds = orc.square_bonds(LX, LY, 1.0)
    _write_deck(d, 16, LX, LY, nod, K, bonds, alg=2, nov=1)
    _run(d)
    m = orc.Model(16, LX, LY, 1, bonds)
    sec = m.build_sector(nod, (orc.wavevector(K[0], LX), orc.wavevector(K[1], LY), 0.0))
    w, V = np.linalg.eigh(m.full_hamiltonian(sec))
    en = _read_cols(os.path.join(d, "out", "energy.dat"))
    assert np.allclose(en[:, 1], w[:3], atol=1e-10)
    # correlations of the ground state are gauge invariant: compare with the oracle
    if abs(w[1] - w[0]) > 1e-8:
        cf = _read_cols(os.path.join(d, "out", "SzSz-corr.dat"))
        pairs = [(int(a), int(b)) for a, b in cf[:, :2]]
        sz_o, sp_o = m.correlations(sec, V[:, 0].astype(np.complex128), pairs)
        assert np.allclose(cf[:, 2], sz_o, atol=1e-8)
        sp = _read_cols(os.path.join(d, "out", "SpSm-corr.dat"))
        assert np.allclose(sp[:, 2] + 1j * sp[:, 3], sp_o, atol=1e-8)


@pytest.mark.gpu
def test_driver_lanczos_chain_with_dsf(tmp_path):
    """ALG=1 on a chain (L=12, N=4, k=0; non-degenerate ground state) with
    S^-(q, w) at q = 2 pi 3/12."""
    print("[gpu] fortran driver ALG=1 + DSF", flush=True)
    d = str(tmp_path)
    L, nod = 12, 4
    bonds = orc.chain_bonds(L, [(1, 1.0, 1.0)])
    q = (2 * np.pi * 3 / L, 0.0, 0.0)
    _write_deck(d, L, L, 1, nod, (0, 0), bonds, alg=1, cal_dsf=1, spsmsz=2, itr_dsf=25, q=q, rfield=0.1, nov=1)
    _run(d)
    m = orc.Model(L, L, 1, 1, bonds)
    sec = m.build_sector(nod, (0.0, 0.0, 0.0))
    H = m.full_hamiltonian(sec)
    w, V = np.linalg.eigh(H)
    ab = _read_cols(os.path.join(d, "out", "alpha_beta_values_E0-tmp.d"))
    new = m.build_sector(nod + 1, (q[0], 0.0, 0.0))
    pos = np.array([[i, 0.0, 0.0] for i in range(L)])
    psi = V[:, 0].astype(np.complex128)
    f = m.sq_a(2, sec, new, np.array(q), pos, psi)
    Hn = m.full_hamiltonian(new)
    sz_new = L / 2.0 - (nod + 1)
    a, b2, n1 = orc.dsf_recurrence(lambda x: Hn @ x - 0.1 * sz_new * x, f, 25)
    assert ab[0, 5] == pytest.approx(n1, rel=1e-9)        # norm1 (gauge invariant)
    assert np.allclose(ab[:10, 3], a[:10], atol=1e-7)       # alpha_n
    assert np.allclose(ab[:10, 4], b2[:10], atol=1e-7)      # beta_n^2


@pytest.mark.gpu
def test_driver_ferromagnetic_state(tmp_path):
    """N_down = 0: THS = 1 and the driver writes only E_FM = sum_b Jz_b / 4
    (main.f90:36-46, calcu_FM_energy, eigen_solver.f90:341-347)."""
    print("[gpu] fortran driver N_down=0", flush=True)
    d = str(tmp_path)
    L = 10
    bonds = orc.chain_bonds(L, [(1, 1.0, 0.8), (2, 0.3, -0.45)])
    _write_deck(d, L, L, 1, 0, (0, 0), bonds, alg=1)
    _run(d)
    en = _read_cols(os.path.join(d, "out", "energy.dat"))
    assert en.shape == (1, 1)
    assert en[0, 0] == pytest.approx(sum(b[3] for b in bonds) / 4.0, abs=1e-13)


@pytest.mark.gpu
@pytest.mark.parametrize("alg", [1, 2])
def test_driver_wavefunction_round_trip(tmp_path, alg):
    """wr_wf = 1 writes the wave function (input_param.f90:119-147, one
    unformatted record per vector); a second run with re_wf = 1 reads it back
    instead of solving (main.f90:66-72, 99-105) and its observables equal the
    first run's, byte for byte in the formatted files."""
    print(f"[gpu] fortran driver wr_wf / re_wf ALG={alg}", flush=True)
    d = str(tmp_path)
    L, nod = 12, 4
    bonds = orc.chain_bonds(L, [(1, 1.0, 0.9)])
    _write_deck(d, L, L, 1, nod, (0, 0), bonds, alg=alg, nov=1, trlan=(1, 3, 12), wr_wf=1, outdir="out1/",
                filewf="wf/")
    _run(d)
    wfs = [f for f in os.listdir(os.path.join(d, "wf")) if f.strip().endswith(".bin")]
    assert len(wfs) == 1 and os.path.exists(os.path.join(d, "wf", "energy.dat"))
    _write_deck(d, L, L, 1, nod, (0, 0), bonds, alg=alg, nov=1, trlan=(1, 3, 12), re_wf=1, outdir="out2/",
                filewf="wf/")
    out = _run(d)
    assert "Eigen solver" not in out  # the solver did not run
    for name in ("local_mag.dat", "SzSz-corr.dat", "SpSm-corr.dat"):
        with open(os.path.join(d, "out1", name)) as a, open(os.path.join(d, "out2", name)) as b:
            assert a.read() == b.read(), name
    # and the vector read back is the ground state: its correlations are the oracle's
    m = orc.Model(L, L, 1, 1, bonds)
    sec = m.build_sector(nod, (0.0, 0.0, 0.0))
    w, V = np.linalg.eigh(m.full_hamiltonian(sec))
    cf = _read_cols(os.path.join(d, "out2", "SzSz-corr.dat"))
    pairs = [(int(a), int(b)) for a, b in cf[:, :2]]
    sz_o, _ = m.correlations(sec, V[:, 0].astype(np.complex128), pairs)
    assert np.allclose(cf[:, 2], sz_o, atol=1e-8)


@pytest.mark.gpu
def test_driver_refuses_incommensurate_q(tmp_path):
    """A DSF momentum transfer q that is not a multiple of 2 pi / L (the
    reference builds an inconsistent target sector at k + q without a check,
    get_expectation_values.f90:430-447) stops the driver with an error that
    names the wave vector."""
    print("[gpu] fortran driver incommensurate q", flush=True)
    d = str(tmp_path)
    L, nod = 12, 3
    bonds = orc.chain_bonds(L, [(1, 1.0, 1.0)])
    q = (2 * np.pi * 2.5 / L, 0.0, 0.0)
    _write_deck(d, L, L, 1, nod, (0, 0), bonds, alg=1, cal_dsf=1, spsmsz=2, itr_dsf=10, q=q, nov=1)
    r = _run(d, expect_ok=False)
    assert r.returncode != 0
    assert "not a multiple of 2 pi" in (r.stdout + r.stderr)


DECKS = {
    # name: (NOS, LX, LY, N_down, K, bonds, ALG, extra deck settings)
    "chain_lanczos_dsf": (12, 12, 1, 4, (0, 0), orc.chain_bonds(12, [(1, 1.0, 1.0)]), 1,
                          dict(cal_dsf=1, spsmsz=2, itr_dsf=25, q=(2 * np.pi * 3 / 12, 0.0, 0.0), rfield=0.1, nov=1)),
    "torus_thick_restart": (16, 4, 4, 3, (1, 0), orc.square_bonds(4, 4, 1.0), 2, dict(nov=1)),
    "chain_j1j2_dense": (10, 10, 1, 3, (2, 0), orc.chain_bonds(10, [(1, 1.0, 0.8), (2, 0.4, 0.4)]), 3, dict(nov=2)),
    "ferromagnet": (10, 10, 1, 0, (0, 0), orc.chain_bonds(10, [(1, 1.0, 0.8)]), 1, dict()),
}


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(REF_DRIVER), reason="reference driver not built (oracle/build_ref.sh)")
@pytest.mark.parametrize("deck", list(DECKS))
def test_reference_driver_runs_on_the_engine(tmp_path, deck):
    """The reference's own main.f90, compiled unchanged against the drop-in
    modules, runs a deck (ALG 1 with the DSF recurrence, 2, 3, and the
    N_down = 0 case) and writes the files the engine's driver writes, with the
    same values: energies to 1e-10, everything else to 1e-8 of each file's
    largest entry."""
    print(f"[gpu] reference main.f90 on the engine: {deck}", flush=True)
    NOS, LX, LY, nod, K, bonds, alg, extra = DECKS[deck]
    outs = {}
    for tag, exe in (("engine", None), ("reference", REF_DRIVER)):
        d = str(tmp_path / tag)
        os.makedirs(d)
        _write_deck(d, NOS, LX, LY, nod, K, bonds, alg=alg, **extra)
        _run(d, exe=exe)
        outs[tag] = d
    fe = sorted(os.listdir(os.path.join(outs["engine"], "out")))
    fr = sorted(os.listdir(os.path.join(outs["reference"], "out")))
    assert fe == fr and len(fe) >= 1, (fe, fr)
    def tokens(path):
        with open(path) as f:
            return [line.split() for line in f if line.strip()]

    def num(t):
        try:
            return float(t.replace("D", "E"))
        except ValueError:
            return None

    for name in fe:
        ta = tokens(os.path.join(outs["engine"], "out", name))
        tb = tokens(os.path.join(outs["reference"], "out", name))
        assert [len(r) for r in ta] == [len(r) for r in tb], name
        xa = [num(t) for r in ta for t in r]
        xb = [num(t) for r in tb for t in r]
        big = max([abs(x) for x in xa if x is not None] + [1.0])
        tol = 1e-10 if name == "energy.dat" else 1e-8 * big
        for ua, ub, sa, sb in zip(xa, xb, [t for r in ta for t in r], [t for r in tb for t in r]):
            if ua is None or ub is None:
                assert sa == sb, (name, sa, sb)  # text (headers, separators)
            else:
                assert abs(ua - ub) <= tol, (name, ua, ub)

