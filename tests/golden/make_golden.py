"""Writes tests/golden/*.npz from the reference's own Fortran modules.

Run in a container that has the reference sources and oracle/_ref built
(oracle/build_ref.sh).  Each fixture holds the inputs (lattice, bonds, wave
vector, a deterministic vector) and the reference's outputs: list_s, list_r,
H.v (ham_to_vec_wave_vector), and where noted the correlation functions
(calcu_cf_trans_2 via mk_list_ij) and S^a_q |v> (calcu_Sq_a) with the target
sector's lists.  Files are plain numpy arrays (load with allow_pickle=False).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import qs3 as orc  # noqa: E402


def vec(dim, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1, 1, dim) + 1j * rng.uniform(-1, 1, dim)


CASES = [
    dict(name="chain12_n4_k3_j1j2", NOS=12, LX=12, LY=1, nod=4, K=(3, 0, 0),
         bonds=orc.chain_bonds(12, [(1, 1.0, 0.7), (2, 0.5, 0.5)]), cf=True, sq=(2, (2 * np.pi * 5 / 12, 0.0, 0.0))),
    dict(name="chain10_n5_k5", NOS=10, LX=10, LY=1, nod=5, K=(5, 0, 0),
         bonds=orc.chain_bonds(10, [(1, 1.0, 1.0)]), cf=False, sq=(1, (np.pi / 5, 0.0, 0.0))),
    dict(name="torus4x3_n3_k11", NOS=12, LX=4, LY=3, nod=3, K=(1, 1, 0),
         bonds=orc.square_bonds(4, 3, 1.0), cf=True, sq=(2, (np.pi / 2, 2 * np.pi / 3, 0.0))),
    dict(name="torus6x6_n3_k12_tri", NOS=36, LX=6, LY=6, nod=3, K=(1, 2, 0),
         bonds=orc.square_bonds(6, 6, 1.0) + [(s, ((s - 1) % 6 + 1) % 6 + 1 + 6 * (((s - 1) // 6 + 1) % 6), 1.0, 1.0)
                                               for s in range(1, 37)], cf=False, sq=(3, (np.pi / 3, 0.0, 0.0))),
]


def main():
    if not orc.RefLib.available():
        sys.exit("oracle/_ref/libqs3ref.so is not built")
    for c in CASES:
        ref = orc.RefLib()
        ref.setup(c["NOS"], c["LX"], c["LY"], 1, c["nod"], c["K"], c["bonds"])
        D = ref.build_sector()
        ls, lr = ref.lists()
        v = vec(D, 12345)
        out = dict(NOS=c["NOS"], LX=c["LX"], LY=c["LY"], nod=c["nod"], K=np.array(c["K"]),
                   bonds=np.array(c["bonds"], dtype=np.float64), list_s=ls, list_r=lr, v=v, hv=ref.hv(v),
                   explist=ref.explist())
        if c["cf"]:
            NOS = c["NOS"]
            pairs = np.array([(a, b) for a in range(1, NOS + 1) for b in range(1, NOS + 1)], dtype=np.int64)
            # per canonical pair (noij > 0): (noij, szsz, spsm) of calcu_cf_trans_2
            cf = []
            for (a, b) in pairs:
                no, sz, sp = ref.cf_pair((b, a), v)   # list_expe_ss(1,:) = column 2
                cf.append((no, sz, sp.real, sp.imag))
            out["cf_pairs"] = pairs
            out["cf_vals"] = np.array(cf)
        spsmsz, q = c["sq"]
        Dn = ref.build_new_sector(spsmsz, q)
        lsn, lrn = ref.new_lists()
        pos = np.array([[i % c["LX"], i // c["LX"], 0.0] for i in range(c["NOS"])], dtype=np.float64)
        out.update(sq_spsmsz=spsmsz, sq_q=np.array(q), sq_pos=pos, sq_list_s=lsn, sq_list_r=lrn, sq_f=ref.sq(pos, v))
        np.savez_compressed(os.path.join(HERE, c["name"] + ".npz"), **out)
        print(c["name"], D, Dn)


if __name__ == "__main__":
    main()
