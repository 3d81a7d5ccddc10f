"""Writes tests/golden/dsflan_*.npz from the reference's own DSFLan program.

DSFLan (source_DSFLan/main.f90) is compiled unmodified by oracle/build_ref.sh
into oracle/_ref/dsflan.exe.  This script writes an alpha/beta file in the
format calcu_DSF_wavevector appends (get_expectation_values.f90:571,600:
ex_i0, dummy, itr, alpha, beta, norm1, E0, sztot per line), runs the program
in a scratch directory, and keeps every STRIDE-th point of its 300,001-point
output (omega = om / 10000, eta = 0.04, as the program fixes them).

Two inputs:
  random   -- 40 levels of seeded random coefficients (beta(1) = 0);
  chain    -- 60 levels of the DSF recurrence of S^-_q |psi0> on a periodic
              Heisenberg chain L = 14 (N_down 3 -> 4, q = 2 pi 3 / 14), from
              the oracle's restatement of calcu_DSF_wavevector.
The coefficients are written with 17 significant digits and the fixture keeps
the values the program read back, so the test evaluates exactly its inputs.
Files are plain numpy arrays (load with allow_pickle=False).
"""
import glob
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import qs3 as orc  # noqa: E402

EXE = os.path.join(ROOT, "oracle", "_ref", "dsflan.exe")
STRIDE = 50


def write_input(path, alpha, beta, norm1, E0):
    with open(path, "w") as f:
        for n in range(len(alpha)):
            f.write(f"0 0 {n + 1} {alpha[n]:.17e} {beta[n]:.17e} {norm1:.17e} {E0:.17e} 0.0\n")


def run_dsflan(alpha, beta, norm1, E0):
    """Runs the reference DSFLan on the coefficients; returns (omega, DSF) of
    all 300,001 points."""
    with tempfile.TemporaryDirectory() as d:
        write_input(os.path.join(d, "alpha_beta_values_E0-tmp.d"), alpha, beta, norm1, E0)
        subprocess.run([EXE], cwd=d, check=True, capture_output=True)
        out = glob.glob(os.path.join(d, "DSF_itrnum*"))
        assert len(out) == 1, out
        data = np.loadtxt(out[0])
    return data[:, 0], data[:, 1]


def chain_coefficients():
    L = 14
    om = orc.Model(L, L, 1, 1, orc.chain_bonds(L, [(1, 1.0, 1.0)]))
    gs = om.build_sector(3, (0.0, 0.0, 0.0))
    H = om.full_hamiltonian(gs)
    w, U = np.linalg.eigh(H)
    psi = U[:, 0]
    q = 2 * np.pi * 3 / L
    tgt = om.build_sector(4, (q, 0.0, 0.0))
    pos = np.array([[i, 0.0, 0.0] for i in range(L)])
    f = om.sq_a(2, gs, tgt, np.array([q, 0.0, 0.0]), pos, psi)
    a, b2, n1 = orc.dsf_recurrence(lambda v: om.hv(tgt, v), f, 60)
    return a, b2, n1, float(w[0])


def main():
    if not os.path.exists(EXE):
        sys.exit("oracle/_ref/dsflan.exe is not built (oracle/build_ref.sh)")
    rng = np.random.default_rng(5)
    a = rng.uniform(-3.0, 1.0, 40)
    b = rng.uniform(0.2, 2.0, 40)
    b[0] = 0.0
    cases = {"random": (a, b, 0.731, -1.2345)}
    ca, cb, cn, cE = chain_coefficients()
    cases["chain"] = (ca, cb, cn, cE)
    for name, (al, be, n1, E0) in cases.items():
        # the values the program reads back from the 17-digit text
        al = np.array([float(f"{x:.17e}") for x in al])
        be = np.array([float(f"{x:.17e}") for x in be])
        n1 = float(f"{n1:.17e}")
        E0 = float(f"{E0:.17e}")
        w, s = run_dsflan(al, be, n1, E0)
        path = os.path.join(HERE, f"dsflan_{name}.npz")
        np.savez(path, alpha=al, beta=be, norm1=np.float64(n1), E0=np.float64(E0), eta=np.float64(0.04),
                 omega=w[::STRIDE], dsf=s[::STRIDE], npoints=np.int64(len(w)))
        print(f"wrote {path}: {len(al)} levels, {len(w[::STRIDE])} of {len(w)} points")


if __name__ == "__main__":
    main()
