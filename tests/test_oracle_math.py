"""Pins the oracle (and its solver restatements) by what mathematics fixes.

* sector dimensions: the character (Burnside) formula, and sum_k dim = C(L,N);
* representatives: brute-force orbit enumeration with the colex order;
* H.v: the union over k of the sector spectra equals the spectrum of the
  N_down block built from bit strings without symmetry; Hermiticity;
* exact spectra: ferromagnet, one magnon, XX free fermions, Bethe ansatz;
* solvers: plain / thick-restart Lanczos against dense eigh;
* DSF: continued fraction against the Lehmann representation, sum rules.
"""
import itertools
import math

import numpy as np
import pytest

from oracle import qs3 as orc


def dim_formula(L, N, K):
    """dim(L,N,k) = (1/L) sum_r e^{-2 pi i k r / L} Fix(T^r) (Burnside), where
    Fix(T^r) = C(g, N g / L) if (L/g) | N, g = gcd(r, L); only k-compatible
    orbits survive, which is what the character sum counts."""
    tot = 0j
    for r in range(L):
        g = math.gcd(r, L)
        per = L // g
        if N % per:
            continue
        tot += np.exp(-2j * np.pi * K * r / L) * math.comb(g, N * g // L)
    v = tot / L
    assert abs(v.imag) < 1e-9
    return int(round(v.real))


@pytest.mark.parametrize("L,N", [(12, 4), (12, 6), (10, 5), (14, 7), (16, 8), (13, 4), (20, 3)])
def test_dimensions_character_formula(L, N):
    m = orc.Model(L, L, 1, 1, orc.chain_bonds(L, [(1, 1.0, 1.0)]))
    tot = 0
    for K in range(L):
        d = m.build_sector(N, m.rk_from_K(K)).dim
        assert d == dim_formula(L, N, K)
        tot += d
    assert tot == math.comb(L, N)


def test_large_dimensions_fast_build():
    # C1: L=36 N=6 k=0 ; C2: L=48 N=6 k=0 (anchored enumeration)
    m = orc.Model(36, 36, 1, 1, orc.chain_bonds(36, [(1, 1.0, 1.0)]))
    assert m.build_sector_chain(6, (0.0, 0.0, 0.0)).dim == 54132 == dim_formula(36, 6, 0)
    m = orc.Model(48, 48, 1, 1, orc.chain_bonds(48, [(1, 1.0, 1.0), (2, 0.5, 0.5)]))
    assert m.build_sector_chain(6, (0.0, 0.0, 0.0)).dim == 255704 == dim_formula(48, 6, 0)


def _brute_reps(L, N, K):
    """Colex-minimal orbit members compatible with k, by enumeration."""
    reps = []
    for comb in itertools.combinations(range(1, L + 1), N):
        orbit = []
        for r in range(1, L + 1):
            img = tuple(sorted(((s - 1 + r) % L) + 1 for s in comb))
            orbit.append((r, img))
        key = lambda c: tuple(reversed(c))  # colex: compare largest element first
        mn = min((img for _, img in orbit), key=key)
        if tuple(comb) != mn:
            continue
        stab = [r for r, img in orbit if img == tuple(comb)]
        coef = sum(np.exp(-2j * np.pi * K * r / L) for r in stab)
        if abs(coef) ** 2 * L / len(stab) > 1e-9:
            reps.append(comb)
    return sorted(reps, key=lambda c: tuple(reversed(c)))  # ascending rank = colex order


@pytest.mark.parametrize("L,N,K", [(9, 3, 0), (9, 3, 3), (10, 4, 5), (8, 4, 2)])
def test_representatives_brute_force(L, N, K):
    m = orc.Model(L, L, 1, 1, orc.chain_bonds(L, [(1, 1.0, 1.0)]))
    sec = m.build_sector(N, m.rk_from_K(K))
    reps = _brute_reps(L, N, K)
    assert [tuple(m.unrank(t, N)) for t in sec.list_s] == reps


def _dense_sz_block(L, N, bonds):
    """N_down block of H from bit strings, no symmetry."""
    confs = list(itertools.combinations(range(1, L + 1), N))
    idx = {c: i for i, c in enumerate(confs)}
    H = np.zeros((len(confs), len(confs)))
    for c, i in idx.items():
        s = set(c)
        for (a, b, jxy, jz) in bonds:
            a, b = int(a), int(b)
            if (a in s) == (b in s):
                H[i, i] += jz / 4
            else:
                H[i, i] -= jz / 4
                t = set(s)
                if a in t:
                    t.remove(a); t.add(b)
                else:
                    t.remove(b); t.add(a)
                H[i, idx[tuple(sorted(t))]] += jxy / 2
    return H


@pytest.mark.parametrize("L,N,fam", [(10, 4, [(1, 1.0, 1.0)]), (9, 3, [(1, 1.0, 0.6), (2, 0.4, -0.3)]),
                                     (8, 4, [(1, 0.7, 1.3)])])
def test_spectra_union_equals_unsymmetrised(L, N, fam):
    bonds = orc.chain_bonds(L, fam)
    m = orc.Model(L, L, 1, 1, bonds)
    evs = []
    for K in range(L):
        sec = m.build_sector(N, m.rk_from_K(K))
        H = m.full_hamiltonian(sec)
        assert np.max(np.abs(H - H.conj().T)) < 1e-13  # Hermitian
        evs.extend(np.linalg.eigvalsh(H))
        rng = np.random.default_rng(K)
        v = rng.normal(size=sec.dim) + 1j * rng.normal(size=sec.dim)
        assert np.allclose(m.hv(sec, v), H @ v, atol=1e-13)
    ref = np.linalg.eigvalsh(_dense_sz_block(L, N, bonds))
    assert np.allclose(np.sort(evs), ref, atol=1e-11)


def test_torus_spectra_union():
    LX, LY, N = 3, 3, 3
    bonds = orc.square_bonds(LX, LY, 1.0)
    m = orc.Model(9, LX, LY, 1, bonds)
    evs = []
    for kx in range(LX):
        for ky in range(LY):
            sec = m.build_sector(N, (orc.wavevector(kx, LX), orc.wavevector(ky, LY), 0.0))
            if sec.dim:
                evs.extend(np.linalg.eigvalsh(m.full_hamiltonian(sec)))
    assert np.allclose(np.sort(evs), np.linalg.eigvalsh(_dense_sz_block(9, N, bonds)), atol=1e-11)


def test_ferromagnet_and_one_magnon():
    L, jxy, jz = 17, 0.9, 1.3
    bonds = orc.chain_bonds(L, [(1, jxy, jz)])
    m = orc.Model(L, L, 1, 1, bonds)
    assert m.fm_energy() == pytest.approx(L * jz / 4)
    for K in range(L):
        sec = m.build_sector(1, m.rk_from_K(K))
        E = np.linalg.eigvalsh(m.full_hamiltonian(sec))[0]
        assert E == pytest.approx(L * jz / 4 - jz + jxy * math.cos(2 * math.pi * K / L), abs=1e-12)
    # FM multiplet member (J < 0): E0 = sum Jz / 4 in every N sector at k = 0
    mf = orc.Model(L, L, 1, 1, orc.chain_bonds(L, [(1, -1.0, -1.0)]))
    for N in (2, 3):
        sec = mf.build_sector(N, (0.0, 0.0, 0.0))
        assert np.linalg.eigvalsh(mf.full_hamiltonian(sec))[0] == pytest.approx(-L / 4, abs=1e-11)


def test_xx_free_fermions():
    """Jz = 0: Jordan-Wigner fermions with hopping Jxy/2; the boundary is
    periodic for odd N and antiperiodic for even N."""
    L = 12
    for N in (3, 4, 5):
        m = orc.Model(L, L, 1, 1, orc.chain_bonds(L, [(1, 1.0, 0.0)]))
        E = min(np.linalg.eigvalsh(m.full_hamiltonian(m.build_sector(N, m.rk_from_K(K))))[0]
                for K in range(L))
        phi = 0.0 if N % 2 else 0.5
        eps = sorted(math.cos(2 * math.pi * (q + phi) / L) for q in range(L))
        assert E == pytest.approx(sum(eps[:N]), abs=1e-11)


def bethe_energy(L, M):
    """XXX ring (J = 1), real rapidities, I_j = -(M-1)/2 .. (M-1)/2:
    L 2 atan(2 x_j) = 2 pi I_j + sum_l 2 atan(x_j - x_l), E = L/4 - sum 2/(4x^2+1)."""
    I = np.arange(M) - (M - 1) / 2.0
    x = 0.5 * np.tan(np.pi * I / L)
    for _ in range(20000):
        rhs = (np.pi * I + np.sum(np.arctan(x[:, None] - x[None, :]), axis=1)) / L
        xn = 0.5 * np.tan(rhs)
        if np.max(np.abs(xn - x)) < 1e-15:
            x = xn
            break
        x = xn
    return L / 4.0 - np.sum(2.0 / (4 * x * x + 1.0))


@pytest.mark.parametrize("L,M", [(10, 2), (12, 4), (14, 4), (16, 6)])
def test_bethe_ansatz_ground_state(L, M):
    m = orc.Model(L, L, 1, 1, orc.chain_bonds(L, [(1, 1.0, 1.0)]))
    sec = m.build_sector(M, (0.0, 0.0, 0.0))
    E = np.linalg.eigvalsh(m.full_hamiltonian(sec))[0]
    assert E == pytest.approx(bethe_energy(L, M), abs=1e-10)


def test_bethe_reference_values():
    """The headline accuracy anchors (L=1000 N=4; L=36 N=6) from the same solver."""
    assert bethe_energy(1000, 4) == pytest.approx(242.000099091345514, abs=1e-9)
    assert bethe_energy(36, 6) == pytest.approx(-2.687328296769200, abs=1e-11)


def test_lanczos_restatements_against_dense():
    L, N = 14, 5
    m = orc.Model(L, L, 1, 1, orc.chain_bonds(L, [(1, 1.0, 0.8), (2, 0.3, 0.3)]))
    sec = m.build_sector(N, m.rk_from_K(3))
    H = m.full_hamiltonian(sec)
    w = np.linalg.eigvalsh(H)
    rng = np.random.default_rng(4)
    v0 = rng.uniform(-1, 1, sec.dim) + 1j * rng.uniform(-1, 1, sec.dim)
    E, psi, it, _, _ = orc.lanczos_plain(lambda v: H @ v, v0, maxitr=sec.dim, conv_tol=1e-14)
    assert E == pytest.approx(w[0], abs=1e-10)
    assert np.linalg.norm(H @ psi - E * psi) < 1e-6
    ene, Y, _ = orc.trlanczos(lambda v: H @ v, v0, NOK=6, NOM=14, conv_tol=1e-14, maxitr=500)
    assert np.allclose(ene[:4], w[:4], atol=1e-10)


def test_continued_fraction_vs_lehmann():
    """The DSF recurrence + continued fraction (exact form) equals the
    Lehmann sum  S(w) = -1/pi Im sum_n |<n|f>|^2 / (w + E0 + i eta - E_n)."""
    L, N = 10, 4
    m = orc.Model(L, L, 1, 1, orc.chain_bonds(L, [(1, 1.0, 1.0)]))
    sec = m.build_sector(N, m.rk_from_K(0))
    H = m.full_hamiltonian(sec)
    w, V = np.linalg.eigh(H)
    rng = np.random.default_rng(5)
    f = rng.normal(size=sec.dim) + 1j * rng.normal(size=sec.dim)
    alpha, beta2, norm1 = orc.dsf_recurrence(lambda v: H @ v, f, sec.dim)
    assert norm1 == pytest.approx(np.vdot(f, f).real)
    assert alpha[0] == pytest.approx((np.vdot(f, H @ f) / np.vdot(f, f)).real)
    E0, eta = w[0], 0.05
    om = np.linspace(0, 6, 301)
    cf = orc.continued_fraction(alpha, beta2, norm1, E0, om, eta, reference_quirk=False)
    wt = np.abs(V.conj().T @ f) ** 2
    leh = -np.imag(np.sum(wt[None, :] / (om[:, None] + E0 + 1j * eta - w[None, :]), axis=1)) / np.pi
    assert np.allclose(cf, leh, atol=1e-8 * np.max(leh))
    # the reference's deepest-level variant differs only through the last level
    q = orc.continued_fraction(alpha, beta2, norm1, E0, om, eta, reference_quirk=True)
    assert np.all(np.isfinite(q))


def test_sq_sum_rule():
    """norm1 = <psi| (S^-_q)^dagger S^-_q |psi> = (1/L) sum_ij e^{iq(Ri-Rj)} <S^+_i S^-_j>
    through the correlation functions (translation invariant chain)."""
    L, N = 10, 3
    bonds = orc.chain_bonds(L, [(1, 1.0, 1.0)])
    m = orc.Model(L, L, 1, 1, bonds)
    sec = m.build_sector(N, (0.0, 0.0, 0.0))
    H = m.full_hamiltonian(sec)
    psi = np.linalg.eigh(H)[1][:, 0].astype(np.complex128)
    pos = np.array([[i, 0.0, 0.0] for i in range(L)])
    for Kq in (0, 3):
        q = np.array([2 * np.pi * Kq / L, 0.0, 0.0])
        new = m.build_sector(N + 1, (q[0], 0.0, 0.0))
        f = m.sq_a(2, sec, new, q, pos, psi)
        pairs = [(a, b) for a in range(1, L + 1) for b in range(1, L + 1)]
        _, spsm = m.correlations(sec, psi, pairs)
        # spsm is indexed like the file lines (a, b) = <S+_b S-_a> (list_expe_ss(1)=b)
        C = np.zeros((L, L), dtype=complex)
        for (a, b), val in zip(pairs, spsm):
            C[b - 1, a - 1] = val
        # <S-_q^dagger S-_q> = (1/L) sum_ij e^{iq(R_i - R_j)} <S+_i S-_j>
        val = np.sum(np.exp(1j * q[0] * (pos[:, 0][:, None] - pos[:, 0][None, :])) * C) / L
        assert np.vdot(f, f).real == pytest.approx(val.real, abs=1e-10)


@pytest.mark.parametrize("L,N,fam", [(10, 4, [(1, 1.0, 1.0)]), (9, 3, [(1, 1.0, 0.6), (2, 0.4, -0.3)]),
                                     (8, 1, [(1, 0.7, 1.3)])])
def test_u1_hv_equals_dense_block(L, N, fam):
    """The oracle's U(1)-only H.v (rows = colex ranks) is the bit-string N_down
    block above, row for row once rank r is mapped to its configuration
    unrank(r): every matrix element, not only the spectrum."""
    bonds = orc.chain_bonds(L, fam)
    m = orc.Model(L, L, 1, 1, bonds)
    Hd = _dense_sz_block(L, N, bonds)
    confs = list(itertools.combinations(range(1, L + 1), N))
    idx = {c: i for i, c in enumerate(confs)}
    D = m.u1_dim(N)
    assert D == len(confs)
    perm = np.array([idx[tuple(int(x) for x in m.unrank(r, N))] for r in range(1, D + 1)])
    assert sorted(perm.tolist()) == list(range(D))  # the ranks enumerate every configuration once
    Hu = np.stack([m.hv_u1(N, np.eye(D)[:, j].astype(complex)) for j in range(D)], axis=1)
    assert np.allclose(Hu, Hd[np.ix_(perm, perm)], atol=1e-14)


def test_incremental_gap_keys_of_chain_hops():
    """The group H.v updates the packed backward-gap keys of a row for a hop
    that moves down spin m by d = +-1 without wrapping (the target empty, so
    the spins keep their order): g[m] -= d and g[m - 1] += d, so key j moves by
    d (unit(j, m - 1) - unit(j, m)), a gap's unit being 2^(kb (N - 1 - t)) for
    its field t = (j - gap) mod N in key j and 0 for the gap the key leaves
    out (t = 0).  Checked against the direct build of the hopped row's keys."""
    import random

    def keys(n, L, N, kb):
        g = [n[q + 1] - n[q] for q in range(N - 1)] + [n[0] + L - n[N - 1]]
        out = []
        for j in range(N):
            k = 0
            for tt in range(1, N):
                k = (k << kb) | g[(j - tt + N) % N]
            out.append(k)
        return out

    def unit(j, gap, N, kb):
        t = (j - gap + N) % N
        return (1 << (kb * (N - 1 - t))) if t else 0

    rnd = random.Random(7)
    n_checked = 0
    for L, N in ((40, 7), (24, 5), (1000, 4), (33, 8), (12, 2), (9, 3)):
        kb = L.bit_length()
        for _ in range(300):
            n = [1] + sorted(rnd.sample(range(2, L + 1), N - 1))
            kb0 = keys(n, L, N, kb)
            for m in range(N):
                for d in (-1, 1):
                    t = n[m] + d
                    if t < 1 or t > L or t in n:
                        continue
                    nn = n[:]
                    nn[m] = t
                    k1 = keys(nn, L, N, kb)
                    for j in range(N):
                        assert kb0[j] + d * (unit(j, (m - 1 + N) % N, N, kb) - unit(j, m, N, kb)) == k1[j]
                        n_checked += 1
    assert n_checked > 50000
