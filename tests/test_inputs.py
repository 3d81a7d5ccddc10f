"""Input tooling (ed_amd/inputs.py) against the reference's generator and the
reference's file formats.

- square_lattice_bonds equals the bond files of source_mk_input_list/
  mk_input_list.f90 (fixture tests/golden/mk_input_list.npz written by the
  program compiled here; live when oracle/_ref/mk_input_list.exe is built);
- the chain and ladder writers produce files the oracle's reader (read_ip's
  list-directed read, input_param.f90:92-103) returns unchanged, with the
  translation symmetry the sectors assume;
- positions follow the example files (input_ex3/list_site_position_1000_type1.dat);
- snap_wavevector snaps q to 2 pi K / L.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from ed_amd import inputs as inp
from oracle import qs3 as orc

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
EXE = os.path.join(ROOT, "oracle", "_ref", "mk_input_list.exe")
GOLD = np.load(os.path.join(HERE, "golden", "mk_input_list.npz"))
KEYS = sorted(k for k in GOLD.files if k.startswith("bonds_"))


@pytest.mark.parametrize("key", KEYS)
def test_square_triangular_bonds_match_reference_generator(key):
    model = int(key.split("_")[1][1])
    lx, ly = (int(x) for x in key.split("_")[2].split("x"))
    jxx, jz = GOLD["couplings" + key[5:]]
    got = np.array(inp.square_lattice_bonds(lx, ly, jxx, jz, model))
    assert np.array_equal(got, GOLD[key])


@pytest.mark.skipif(not os.path.exists(EXE), reason="reference generator not built (oracle/build_ref.sh)")
def test_square_bonds_match_live_reference_generator(tmp_path):
    nml = '&list_input\nmodel=2, lx=7, ly=5, jxx=0.25d0, jz=-1.5d0, filexxz="b.dat"\n/\n'
    subprocess.run([EXE], input=nml, text=True, cwd=tmp_path, check=True, capture_output=True)
    ref = np.loadtxt(tmp_path / "b.dat")
    assert np.array_equal(np.array(inp.square_lattice_bonds(7, 5, 0.25, -1.5, 2)), ref)


def _bond_set(bonds):
    return sorted((min(int(i), int(j)), max(int(i), int(j)), float(a), float(b)) for i, j, a, b in bonds)


def test_square_lattice_bonds_are_the_nearest_neighbours():
    # every site of the lx x ly square lattice has 4 distinct neighbours (lx, ly >= 3)
    for lx, ly in ((4, 4), (5, 3)):
        b = inp.square_lattice_bonds(lx, ly, 1.0, 1.0, 1)
        assert len(b) == 2 * lx * ly
        deg = np.zeros(lx * ly + 1, dtype=int)
        for i, j, _, _ in b:
            deg[i] += 1
            deg[j] += 1
        assert np.all(deg[1:] == 4)
        assert len(set((i, j) for i, j, _, _ in b)) == len(b)


@pytest.mark.parametrize("fam", [[(1, 1.0, 1.0)], [(1, 1.0, 0.7), (2, 0.5, 0.25)]])
def test_chain_writer_round_trip(tmp_path, fam):
    L = 12
    bonds = inp.chain_bonds(L, fam)
    assert bonds == [tuple(b) for b in orc.chain_bonds(L, fam)]
    p = tmp_path / "list_xxz_term_12.dat"
    inp.write_bond_list(p, bonds)
    back = orc.read_bond_file(str(p), len(bonds))
    assert np.array_equal(np.asarray(back, dtype=float), np.asarray(bonds, dtype=float))


def test_ladder_bonds_and_translations():
    L = 6
    b = inp.ladder_bonds(L, 1.0, 0.8, 0.5, 0.4)
    assert len(b) == 3 * L
    deg = np.zeros(2 * L + 1, dtype=int)
    for i, j, _, _ in b:
        deg[i] += 1
        deg[j] += 1
    assert np.all(deg[1:] == 3)
    # the bond set maps to itself under x and y translations (state_lists.f90:36-64)
    m = orc.Model(2 * L, L, 2, 1, b)
    sx, sy, _ = m.shift_tables()
    for sh in (sx, sy):
        moved = [(sh[i - 1], sh[j - 1], a, z) for i, j, a, z in b]
        assert _bond_set(moved) == _bond_set(b)


def test_ladder_spectrum_is_the_union_over_k():
    """The ladder's (KX, KY) sectors together hold the unsymmetrised block."""
    L, N = 4, 3
    b = inp.ladder_bonds(L, 1.0, 1.0, 0.7, 0.7)
    m = orc.Model(2 * L, L, 2, 1, b)
    ev = []
    for kx in range(L):
        for ky in range(2):
            sec = m.build_sector(N, m.rk_from_K(kx, ky))
            if sec.dim:
                ev.extend(np.linalg.eigvalsh(m.full_hamiltonian(sec)))
    u1 = orc.Model(2 * L, 2 * L, 1, 1, b, ncol=N + 1)
    full = np.linalg.eigvalsh(np.array([u1.hv_u1(N, e) for e in np.eye(math.comb(2 * L, N))]).T)
    assert np.allclose(np.sort(ev), full, atol=1e-10)


def test_positions_writer(tmp_path):
    p = tmp_path / "pos.dat"
    inp.write_positions(p, 4, 3, 2)
    a = np.loadtxt(p)
    assert a.shape == (24, 3)
    # site = x + LX (y - 1) + LX LY (z - 1): site 6 is (x, y, z) = (2, 2, 1)
    assert np.array_equal(a[5], [1.0, 1.0, 0.0])
    with open(p) as f:
        assert f.readline() == "  0.000000000000000E+00  0.000000000000000E+00  0.000000000000000E+00\n"


def test_input_dat_writer(tmp_path):
    p = tmp_path / "input.dat"
    inp.write_input_dat(p, input_parameters={"NOS": 12, "NOD": 4, "FILExxz": "b.dat"},
                        input_lancz={"lnc_ene_conv": 1e-14, "maxitr": 100})
    txt = p.read_text()
    assert txt.index("&input_parameters") < txt.index("&input_lancz")
    assert 'FILExxz = "b.dat",' in txt and "lnc_ene_conv = 1d-14," in txt and "NOS = 12," in txt


def test_snap_wavevector():
    K, rk, dev = inp.snap_wavevector(math.pi / 2, 10)  # input_ex3's QZ on LZ = 10: K = 2.5
    assert dev == pytest.approx(0.5) and K in (2, 3)
    K, rk, dev = inp.snap_wavevector(2 * math.pi * 16 / 1000, 1000)
    assert K == 16 and dev < 1e-12 and rk == orc.wavevector(16, 1000)
    K, rk, dev = inp.snap_wavevector(2 * math.pi / 64, 1000)  # 15.625
    assert K == 16 and dev == pytest.approx(0.375)
