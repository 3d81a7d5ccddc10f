"""Writes tests/golden/mk_input_list.npz from the reference's own bond-list
generator (source_mk_input_list/mk_input_list.f90, compiled unmodified by
oracle/build_ref.sh into oracle/_ref/mk_input_list.exe).  For each (model,
lx, ly) the program reads its namelist from stdin and writes the bond file;
the fixture keeps the bonds (i, j, Jxy, Jz) as arrays named
bonds_m{model}_{lx}x{ly}.  Load with allow_pickle=False."""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(ROOT, "oracle", "_ref", "mk_input_list.exe")
CASES = [(1, 4, 4, 1.0, 1.0), (1, 6, 3, 1.0, 0.5), (1, 3, 5, -1.0, -1.0), (2, 4, 3, 1.0, 0.5), (2, 6, 6, 1.0, 1.0),
         (2, 5, 4, 0.7, 1.3)]


def run(model, lx, ly, jxx, jz):
    with tempfile.TemporaryDirectory() as d:
        nml = f'&list_input\nmodel={model}, lx={lx}, ly={ly}, jxx={jxx!r}d0, jz={jz!r}d0, filexxz="b.dat"\n/\n'
        subprocess.run([EXE], input=nml, text=True, cwd=d, check=True, capture_output=True)
        return np.loadtxt(os.path.join(d, "b.dat"), ndmin=2)


def main():
    if not os.path.exists(EXE):
        sys.exit("oracle/_ref/mk_input_list.exe is not built (oracle/build_ref.sh)")
    arrays = {}
    for model, lx, ly, jxx, jz in CASES:
        arrays[f"bonds_m{model}_{lx}x{ly}"] = run(model, lx, ly, jxx, jz)
        arrays[f"couplings_m{model}_{lx}x{ly}"] = np.array([jxx, jz])
    path = os.path.join(HERE, "mk_input_list.npz")
    np.savez(path, **arrays)
    print(f"wrote {path}: {len(CASES)} lattices")


if __name__ == "__main__":
    main()
