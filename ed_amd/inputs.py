"""Input files of QS3 for chains, ladders and square / triangular lattices.

The QS3 driver reads a bond list (list_xxz_term_*.dat: lines "i j Jxy Jz",
read list-directed by read_ip, input_param.f90:92-103), site positions
(list_site_position_*.dat: "x y z" per site, get_expectation_values.f90:30-35)
and the namelists of input.dat (input_param.f90:63-73, lanczos.f90:28,
eigen_solver.f90:26).  This module writes them:

  square_lattice_bonds  -- the bond enumeration of the reference's generator
                           (source_mk_input_list/mk_input_list.f90:1-71):
                           model 1 square, model 2 triangular, lx by ly, periodic
  chain_bonds           -- periodic chain, any set of bond distances
  ladder_bonds          -- periodic two-leg ladder (LX = L, LY = 2)
  write_bond_list, write_positions, write_input_dat
  snap_wavevector       -- nearest wave vector 2 pi K / L the lattice carries

A transferred wave vector q of the DSF (input_dynamic QX, QY, QZ) must be a
multiple of 2 pi / L_a in every direction: the target sector is built at
k + q (get_expectation_values.f90:430-447).  snap_wavevector gives the nearest
such q; the engine refuses any other (qs3x_sector_build).
"""
import math

__all__ = ["square_lattice_bonds", "chain_bonds", "ladder_bonds", "site_positions", "write_bond_list",
           "write_positions", "write_input_dat", "snap_wavevector"]


def square_lattice_bonds(lx, ly, jxx, jz, model=1):
    """Bonds (i, j, Jxy, Jz) in the order mk_input_list.f90 writes them:
    bond_x, bond_y and, for the triangular lattice (model 2), bond_diag.
    Sites p = 1 .. lx ly run over y fastest (p = y + ly (x - 1)); every pair is
    written as (min, max)."""
    if model not in (1, 2):
        raise ValueError("model must be 1 (square) or 2 (triangular)")
    out = []

    def bond(a, b):
        out.append((min(a, b), max(a, b), float(jxx), float(jz)))

    # bond_x: site p to p + ly, then the wrap of the last column to the first
    p = 1
    for _ in range(lx - 1):
        for _ in range(ly):
            bond(p, p + ly)
            p += 1
    pd = 1
    for _ in range(ly):
        bond(p, pd)
        p += 1
        pd += 1
    # bond_y: p to p + 1 inside a column, then the column's wrap
    p = 1
    for _ in range(lx):
        for _ in range(ly - 1):
            bond(p, p + 1)
            p += 1
        bond(p, p - ly + 1)
        p += 1
    if model == 2:
        # bond_diag: p to p + 2 ly - 1 at the column's first site, p + ly - 1
        # at the others; then the last column wraps to the first
        p = 1
        for _ in range(lx - 1):
            bond(p, p + 2 * ly - 1)
            p += 1
            for _ in range(2, ly + 1):
                bond(p, p + ly - 1)
                p += 1
        pd = -ly + 1
        bond(p, pd + 2 * ly - 1)
        p += 1
        pd += 1
        for _ in range(2, ly + 1):
            bond(p, pd + ly - 1)
            p += 1
            pd += 1
    return out


def chain_bonds(L, couplings):
    """Periodic chain of L sites: for each (d, Jxy, Jz) in couplings, bonds
    (i, i + d mod L) for i = 1 .. L."""
    out = []
    for d, jxy, jz in couplings:
        for i in range(1, L + 1):
            out.append((i, (i - 1 + d) % L + 1, float(jxy), float(jz)))
    return out


def ladder_bonds(L, j_leg, jz_leg, j_rung, jz_rung):
    """Periodic two-leg ladder on the LX = L, LY = 2 lattice (site = x + L (y -
    1), QS3's numbering, state_lists.f90:35): the legs (x, x + 1) of both rows,
    then one rung (x, x + L) per x.  The bond set maps to itself under the
    x and y translations, so the (KX, KY) sectors with KY = 0, 1 apply."""
    if L < 3:
        raise ValueError("a periodic ladder needs L >= 3")
    out = []
    for y in (1, 2):
        for x in range(1, L + 1):
            out.append((x + L * (y - 1), x % L + 1 + L * (y - 1), float(j_leg), float(jz_leg)))
    for x in range(1, L + 1):
        out.append((x, x + L, float(j_rung), float(jz_rung)))
    return out


def site_positions(LX, LY=1, LZ=1):
    """(x - 1, y - 1, z - 1) for sites 1 .. LX LY LZ in QS3's order (x fastest)."""
    return [(float(x), float(y), float(z)) for z in range(LZ) for y in range(LY) for x in range(LX)]


def write_bond_list(path, bonds):
    """list_xxz_term_*.dat as the examples hold it: '(2i5, 2es23.15)' lines."""
    with open(path, "w") as f:
        for i, j, jxy, jz in bonds:
            f.write(f"{int(i):5d}{int(j):5d}{jxy:23.15E}{jz:23.15E}\n")


def write_positions(path, LX, LY=1, LZ=1):
    """list_site_position_*.dat: 'x y z' per site as '(3es23.15)'."""
    with open(path, "w") as f:
        for x, y, z in site_positions(LX, LY, LZ):
            f.write(f"{x:23.15E}{y:23.15E}{z:23.15E}\n")


def _nl_value(v):
    if isinstance(v, bool):
        return ".true." if v else ".false."
    if isinstance(v, int):
        return str(v)
    if isinstance(v, float):
        return repr(v).replace("e", "d") if "e" in repr(v) else repr(v) + "d0"
    return '"' + str(v) + '"'


NAMELISTS = ("input_parameters", "input_static", "input_dynamic", "input_lancz", "input_TRLan")


def write_input_dat(path, **groups):
    """input.dat: the namelists in the order read_ip reads them, each from a
    dict of name -> value (int, float, str); missing groups are omitted."""
    with open(path, "w") as f:
        for g in NAMELISTS:
            if g not in groups:
                continue
            f.write(f"&{g}\n")
            for k, v in groups[g].items():
                f.write(f"{k} = {_nl_value(v)},\n")
            f.write("&end\n\n")


def snap_wavevector(q, L):
    """Nearest wave vector the L translations carry: (K, 2 pi K / L as
    read_ip evaluates it, |q L / (2 pi) - K|).  A deviation above ~1e-8 means
    q itself would be refused as an incommensurate momentum transfer."""
    PI = math.acos(-1.0)
    x = float(q) * L / (2.0 * PI)
    K = int(math.floor(x + 0.5))
    return K, 2.0 * PI * float(K) / float(L), abs(x - K)
