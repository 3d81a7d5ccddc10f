"""Golden vectors written by the reference's own modules (tests/golden/
make_golden.py): sector lists, one H.v, correlation functions and S^a_q.
CPU: the oracle reproduces them.  GPU: the engine reproduces them."""
import glob
import os

import numpy as np
import pytest

from oracle import qs3 as orc

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = sorted(f for f in glob.glob(os.path.join(HERE, "golden", "*.npz"))
               if not os.path.basename(f).startswith(("dsflan_", "mk_input_list")))  # test_dsflan_reference.py, test_inputs.py
IDS = [os.path.basename(f)[:-4] for f in FILES]


def _load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _rk(g):
    K = g["K"]
    return (orc.wavevector(K[0], int(g["LX"])), orc.wavevector(K[1], int(g["LY"])), 0.0)


@pytest.mark.parametrize("path", FILES, ids=IDS)
def test_oracle_reproduces_golden(path):
    g = _load(path)
    NOS, LX, LY, nod = int(g["NOS"]), int(g["LX"]), int(g["LY"]), int(g["nod"])
    m = orc.Model(NOS, LX, LY, 1, g["bonds"])
    sec = m.build_sector(nod, _rk(g))
    assert np.array_equal(sec.list_s, g["list_s"]) and np.array_equal(sec.list_r, g["list_r"])
    assert np.array_equal(sec.explist, g["explist"])
    hv = m.hv(sec, g["v"])
    assert np.max(np.abs(hv - g["hv"])) <= 1e-14 * np.max(np.abs(g["hv"]))
    if "cf_vals" in g:
        scale = np.vdot(g["v"], g["v"]).real
        for (a, b), (no, sz, spr, spi) in zip(g["cf_pairs"], g["cf_vals"]):
            n2, s2, p2 = m.cf_pair(sec, g["v"], (int(b), int(a)))
            assert n2 == int(no)
            if n2:
                assert abs(s2 - sz) <= 1e-13 * scale and abs(p2 - (spr + 1j * spi)) <= 1e-13 * scale
    spsmsz, q = int(g["sq_spsmsz"]), g["sq_q"]
    dn = {1: -1, 2: 1, 3: 0}[spsmsz]
    rk = _rk(g)
    new = m.build_sector(nod + dn, (rk[0] + q[0], rk[1] + q[1], rk[2] + q[2]))
    assert np.array_equal(new.list_s, g["sq_list_s"]) and np.array_equal(new.list_r, g["sq_list_r"])
    f = m.sq_a(spsmsz, sec, new, q, g["sq_pos"], g["v"])
    assert np.max(np.abs(f - g["sq_f"])) <= 1e-13 * max(1.0, np.max(np.abs(g["sq_f"])))


@pytest.mark.gpu
@pytest.mark.parametrize("path", FILES, ids=IDS)
def test_engine_reproduces_golden(path):
    print(f"[gpu] golden {os.path.basename(path)}", flush=True)
    from ed_amd import qs3 as q3
    g = _load(path)
    NOS, LX, LY, nod = int(g["NOS"]), int(g["LX"]), int(g["LY"]), int(g["nod"])
    gm = q3.Model(NOS, LX, LY, 1, g["bonds"])
    rk = _rk(g)
    sec = q3.Sector.build(gm, nod, rk)
    ls, lr = sec.lists()
    assert np.array_equal(ls, g["list_s"]) and np.array_equal(lr, g["list_r"])
    hv = sec.ham_to_vec_wave_vector(g["v"])
    assert np.max(np.abs(hv - g["hv"])) <= 1e-12 * np.max(np.abs(g["hv"]))
    if "cf_vals" in g:
        pairs = g["cf_pairs"]
        sz, sp = q3.get_cf_2_wave_vector(sec, g["v"], pairs)
        # expand the reference's per-orbit values to every pair of the orbit
        m = orc.Model(NOS, LX, LY, 1, g["bonds"])
        full_sz = {}
        full_sp = {}
        for (a, b), (no, s, pr, pi) in zip(pairs, g["cf_vals"]):
            if int(no) == 0:
                continue
            for (x, y) in m.mk_list_ij((int(b), int(a))):
                full_sz[(x, y)] = s
                full_sp[(x, y)] = pr + 1j * pi
        scale = np.vdot(g["v"], g["v"]).real
        for j, (a, b) in enumerate(pairs):
            key = (int(b), int(a))
            assert abs(sz[j] - full_sz[key]) <= 1e-12 * scale
            assert abs(sp[j] - full_sp[key]) <= 1e-12 * scale
    spsmsz, q = int(g["sq_spsmsz"]), g["sq_q"]
    dn = {1: -1, 2: 1, 3: 0}[spsmsz]
    new = q3.Sector.build(gm, nod + dn, (rk[0] + q[0], rk[1] + q[1], rk[2] + q[2]))
    ls2, lr2 = new.lists()
    assert np.array_equal(ls2, g["sq_list_s"]) and np.array_equal(lr2, g["sq_list_r"])
    f = q3.calcu_Sq_a(spsmsz, sec, new, q, g["sq_pos"], g["v"])
    assert np.max(np.abs(f - g["sq_f"])) <= 1e-12 * max(1.0, np.max(np.abs(g["sq_f"])))
