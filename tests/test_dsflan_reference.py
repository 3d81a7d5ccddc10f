"""DSF_analysis (source_DSFLan/main.f90:40-94) against the reference program.

The fixtures tests/golden/dsflan_*.npz hold the output of the reference's own
DSFLan, compiled unmodified (oracle/_ref/dsflan.exe, tests/golden/
make_dsflan_golden.py), including its deepest-level evaluation (main.f90:61-65:
the imaginary part of the last level is not divided by the denominator).
Both the oracle's continued_fraction and the engine's qs3x_continued_fraction
must reproduce it to 1e-13 (relative to the largest |S|); the exact fraction
must not, so the fixture tells the two forms apart.  When the reference
program is built here, the full 300,001-point output is compared too.
"""
import glob
import os
import subprocess

import numpy as np
import pytest

from oracle import qs3 as orc

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
EXE = os.path.join(ROOT, "oracle", "_ref", "dsflan.exe")
CASES = ["random", "chain"]


def load(name):
    d = np.load(os.path.join(HERE, "golden", f"dsflan_{name}.npz"))
    stride = (int(d["npoints"]) - 1) // (len(d["omega"]) - 1)
    omega = np.arange(0, int(d["npoints"]), stride, dtype=np.float64) / 10000.0
    assert np.allclose(omega, d["omega"], rtol=0, atol=1e-15)
    return d, omega


def relerr(got, ref):
    return np.max(np.abs(got - ref)) / np.max(np.abs(ref))


@pytest.mark.parametrize("name", CASES)
def test_oracle_quirk_matches_reference_dsflan(name):
    d, omega = load(name)
    s = orc.continued_fraction(d["alpha"], d["beta"], float(d["norm1"]), float(d["E0"]), omega,
                               float(d["eta"]), reference_quirk=True)
    assert relerr(s, d["dsf"]) < 1e-13
    # the exact fraction differs from the reference's deepest level
    s_exact = orc.continued_fraction(d["alpha"], d["beta"], float(d["norm1"]), float(d["E0"]), omega,
                                     float(d["eta"]), reference_quirk=False)
    assert relerr(s_exact, d["dsf"]) > 1e-10


@pytest.mark.parametrize("name", CASES)
def test_engine_quirk_matches_reference_dsflan(name):
    from ed_amd import qs3 as q3
    d, omega = load(name)
    s = q3.dsf_analysis(d["alpha"], d["beta"], float(d["norm1"]), float(d["E0"]), omega, float(d["eta"]),
                        reference_quirk=True)
    assert relerr(s, d["dsf"]) < 1e-13


def test_exact_fraction_matches_lehmann_sum():
    """The exact form (reference_quirk=False) is the resolvent of the Jacobi
    matrix: norm1 * sum_j |U_0j|^2 eta / pi / ((w + E0 - e_j)^2 + eta^2)."""
    d, omega = load("chain")
    a, b = d["alpha"], d["beta"]
    M = len(a)
    T = np.diag(a) + np.diag(np.sqrt(b[1:M]), 1) + np.diag(np.sqrt(b[1:M]), -1)
    e, U = np.linalg.eigh(T)
    eta, E0, n1 = float(d["eta"]), float(d["E0"]), float(d["norm1"])
    leh = n1 * np.sum(U[0] ** 2 * eta / np.pi / ((omega[:, None] + E0 - e[None, :]) ** 2 + eta ** 2), axis=1)
    s = orc.continued_fraction(a, b, n1, E0, omega, eta, reference_quirk=False)
    assert relerr(s, leh) < 1e-11


@pytest.mark.skipif(not os.path.exists(EXE), reason="reference DSFLan not built (oracle/build_ref.sh)")
def test_oracle_quirk_matches_live_dsflan_all_points(tmp_path):
    d, _ = load("chain")
    with open(tmp_path / "alpha_beta_values_E0-tmp.d", "w") as f:
        for n in range(len(d["alpha"])):
            f.write(f"0 0 {n + 1} {d['alpha'][n]:.17e} {d['beta'][n]:.17e} {float(d['norm1']):.17e} "
                    f"{float(d['E0']):.17e} 0.0\n")
    subprocess.run([EXE], cwd=tmp_path, check=True, capture_output=True)
    out = glob.glob(str(tmp_path / "DSF_itrnum*"))
    assert len(out) == 1
    data = np.loadtxt(out[0])
    assert data.shape == (300001, 2)
    omega = np.arange(300001, dtype=np.float64) / 10000.0
    s = orc.continued_fraction(d["alpha"], d["beta"], float(d["norm1"]), float(d["E0"]), omega,
                               float(d["eta"]), reference_quirk=True)
    assert relerr(s, data[:, 1]) < 1e-13
