"""GPU parity of the batched multi-q DSF path against the per-sector path.

qs3x_dsf_sminus_batch builds S^-_q |psi> for many wave numbers K into one
all-k sector and runs the DSF recurrence (calcu_DSF_wavevector,
get_expectation_values.f90:379-605) on all of them as one block of
interleaved vectors.  Its outputs must equal, q by q, those of the per-sector
path: qs3x_apply_sq into the k = 2 pi K / L sector (pinned against the oracle
in test_gpu_solvers.py) followed by qs3x_dsf_recurrence.  Tolerances: norm1
1e-12 relative; alpha, beta^2 1e-9 (FP64) / 1e-4 (FP32) relative to their
scale over the first steps (late coefficients of a recurrence without
reorthogonalisation amplify rounding differences in any precision)."""
import numpy as np
import pytest

from oracle import qs3 as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _q3():
    from ed_amd import qs3 as q3
    return q3


def dimer(L, j1, j2, jz1, jz2):
    return [(i, i % L + 1, j1 if i % 2 else j2, jz1 if i % 2 else jz2) for i in range(1, L + 1)]


def test_all_k_sector_is_the_k0_list():
    """Every representative has ro > 0 at k = 0, so the all-k sector's rows are
    the k = 0 sector's list_s (state_lists.f90:225-292)."""
    print("[gpu] all-k sector", flush=True)
    q3 = _q3()
    for L, N in ((12, 3), (16, 4), (15, 5), (20, 2)):
        bonds = orc.chain_bonds(L, [(1, 1.0, 1.0)])
        om = orc.Model(L, L, 1, 1, bonds)
        o0 = om.build_sector(N, (0.0, 0.0, 0.0))
        gm = q3.Model(L, L, 1, 1, bonds)
        ak = q3.Sector.build_all_k(gm, N)
        ls, _ = ak.lists()
        assert np.array_equal(ls, o0.list_s)
        # the union of the k sectors is the same set
        union = set()
        for K in range(L):
            union |= set(om.build_sector(N, om.rk_from_K(K)).list_s.tolist())
        assert sorted(union) == ls.tolist()


CASES = [
    # name, L, N_gs, bonds, K list, steps
    ("chain12_n2", 12, 2, orc.chain_bonds(12, [(1, 1.0, 1.0)]), list(range(12)), 25),
    ("chain16_n3_xxz", 16, 3, orc.chain_bonds(16, [(1, 1.0, 0.7)]), [0, 1, 4, 8, 8, 15, 3], 25),
    ("chain14_n3_dimer", 14, 3, dimer(14, 1.0, 0.6, 0.9, 0.4), list(range(14)), 25),
    ("chain18_n5", 18, 5, orc.chain_bonds(18, [(1, 1.0, 1.0)]), [0, 3, 6, 9, 2], 20),
]


@pytest.mark.parametrize("prec_name", ["fp64", "fp32"])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_dsf_batch_matches_per_sector(case, prec_name):
    name, L, Ngs, bonds, Ks, steps = case
    print(f"[gpu] dsf batch {name} {prec_name}", flush=True)
    q3 = _q3()
    prec = q3.FP64 if prec_name == "fp64" else q3.FP32
    gm = q3.Model(L, L, 1, 1, bonds)
    gsec = q3.Sector.build(gm, Ngs, (0.0, 0.0, 0.0))
    rng = np.random.default_rng(11)
    psi = rng.uniform(-1, 1, gsec.dim) + 1j * rng.uniform(-1, 1, gsec.dim)
    pos = np.array([[i, 0.0, 0.0] for i in range(L)], dtype=float)
    rfield = 0.2
    a_b, b_b, n_b, st = q3.dsf_sminus_batch(gsec, Ks, pos, psi, steps, rfield=rfield, prec=prec)
    assert st["steps"] >= 1
    tol = 1e-9 if prec == q3.FP64 else 1e-4
    nchk = 12
    for j, K in enumerate(Ks):
        rk = q3.wavevector(K, L)
        tgt = q3.Sector.build(gm, Ngs + 1, (rk, 0.0, 0.0))
        if tgt.dim == 0:
            assert n_b[j] == 0.0
            continue
        f = q3.calcu_Sq_a(2, gsec, tgt, np.array([rk, 0.0, 0.0]), pos, psi)
        n1 = np.vdot(f, f).real
        assert n_b[j] == pytest.approx(n1, rel=1e-12, abs=1e-300)
        if n1 < 1e-20:
            continue
        s = min(steps, tgt.dim)
        a_s, b_s, n_s = q3.dsf_recurrence(tgt, f, s, rfield=rfield, prec=prec)
        # a recurrence longer than half the sector dimension has resolved the
        # spectrum and loses orthogonality (FP32 first)
        k = min(nchk, s, max(2, tgt.dim // 2))
        # an invariant Krylov subspace (e.g. K = 0 of an isotropic chain, where
        # S^- commutes with H) ends the recurrence: later coefficients are noise
        for i in range(1, k):
            if b_s[i] < 1e-10 * max(np.max(b_s[1:i + 1]), 1e-300) or b_s[i] < 1e-24:
                k = i
                break
        sa = max(1.0, np.max(np.abs(a_s[:k])))
        sb = max(1.0, np.max(np.abs(b_s[:k])))
        assert np.max(np.abs(a_b[j, :k] - a_s[:k])) <= tol * sa, (K, a_b[j, :k], a_s[:k])
        assert np.max(np.abs(b_b[j, :k] - b_s[:k])) <= tol * sb, (K, b_b[j, :k], b_s[:k])


def test_dsf_batch_fp32_vs_fp64():
    """FP32 block against FP64 block on the same inputs: early coefficients to
    FP32 resolution of the shifted operator."""
    print("[gpu] dsf batch fp32 vs fp64", flush=True)
    q3 = _q3()
    L, Ngs = 20, 3
    bonds = orc.chain_bonds(L, [(1, 1.0, 1.0)])
    gm = q3.Model(L, L, 1, 1, bonds)
    gsec = q3.Sector.build(gm, Ngs, (0.0, 0.0, 0.0))
    E0, psi, _ = q3.lanczos_routines(gsec, maxitr=400, lnc_ene_conv=1e-13, prec=q3.FP64, seed=3)
    pos = np.array([[i, 0.0, 0.0] for i in range(L)], dtype=float)
    Ks = list(range(1, L))
    a64, b64, n64, _ = q3.dsf_sminus_batch(gsec, Ks, pos, psi, 30, prec=q3.FP64)
    a32, b32, n32, _ = q3.dsf_sminus_batch(gsec, Ks, pos, psi, 30, prec=q3.FP32)
    assert np.allclose(n32, n64, rtol=1e-12)
    assert np.max(np.abs(a32[:, :10] - a64[:, :10])) <= 1e-4 * max(1.0, np.max(np.abs(a64[:, :10])))
    assert np.max(np.abs(b32[:, :10] - b64[:, :10])) <= 1e-4 * max(1.0, np.max(np.abs(b64[:, :10])))


@pytest.mark.parametrize("prec_name,graph", [("fp64", True), ("fp32", True), ("fp64", False)])
def test_lanczos_group_matches_per_sector(prec_name, graph):
    """All (N_down, k) sectors of an XXZ chain in one group: each sector's E0
    equals the per-sector plain Lanczos (same seed, same recurrence) and, in
    FP64, the dense ground-state energy of the oracle's sector Hamiltonian."""
    print(f"[gpu] lanczos group {prec_name} graph={graph}", flush=True)
    q3 = _q3()
    prec = q3.FP64 if prec_name == "fp64" else q3.FP32
    L = 12
    bonds = orc.chain_bonds(L, [(1, 1.0, 0.8)])
    gm = q3.Model(L, L, 1, 1, bonds)
    om = orc.Model(L, L, 1, 1, bonds)
    secs, keys = [], []
    for N in range(1, 7):
        for K in range(L):
            s = q3.Sector.build(gm, N, (q3.wavevector(K, L), 0.0, 0.0))
            if s.dim:
                secs.append(s)
                keys.append((N, K))
    conv = 1e-12
    E, its, st = q3.lanczos_group(secs, maxitr=500, minitr=10, itrint=5, lnc_ene_conv=conv, prec=prec, seed=5,
                                  use_graph=graph)
    assert st["groups"] == 6
    # FP32: the group kernel and the per-sector kernel round differently, and
    # each FP32 recurrence ends within about 3e-6 of the FP64 value on these
    # sectors (measured up to 2.8e-6), so both are held to 5e-6 of the FP64
    # per-sector E0 and to each other.
    tol = 1e-9 if prec == q3.FP64 else 5e-6
    for (N, K), s, e in zip(keys, secs, E):
        e1, _, _ = q3.lanczos_routines(s, maxitr=500, minitr=10, itrint=5, lnc_ene_conv=conv, prec=prec, seed=5,
                                       want_vector=False)
        assert abs(e - e1) <= tol * max(1.0, abs(e1)), (N, K, e, e1)
        if prec == q3.FP32:
            e64, _, _ = q3.lanczos_routines(s, maxitr=500, minitr=10, itrint=5, lnc_ene_conv=conv, prec=q3.FP64,
                                            seed=5, want_vector=False)
            assert abs(e - e64) <= tol * max(1.0, abs(e64)), (N, K, e, e64)
            assert abs(e1 - e64) <= tol * max(1.0, abs(e64)), (N, K, e1, e64)
        if prec == q3.FP64:
            osec = om.build_sector(N, om.rk_from_K(K))
            H = om.full_hamiltonian(osec)
            ed = np.linalg.eigvalsh(H)[0]
            assert abs(e - ed) <= 1e-9 * max(1.0, abs(ed)), (N, K, e, ed)


@pytest.mark.parametrize("L,J2,prec_name,graph", [(24, 0.0, "fp64", True), (24, 0.0, "fp32", False),
                                                  (18, 0.4, "fp32", True), (18, 0.4, "fp64", False)])
def test_lanczos_group_hop_cache_bit_identical(L, J2, prec_name, graph):
    """Option group_hop_cache: the group's first step records every row's
    hops and the later steps read them instead of searching for
    representatives again.  The arithmetic is the same, so energies and step
    counts equal those without the cache to the bit, on nearest-neighbour
    chains (2 hop slots per spin) and J1-J2 chains (4), in both precisions,
    with and without the captured step."""
    print(f"[gpu] lanczos group hop cache L={L} J2={J2} {prec_name} graph={graph}", flush=True)
    q3 = _q3()
    prec = q3.FP64 if prec_name == "fp64" else q3.FP32
    terms = [(1, 1.0, 0.8)] + ([(2, J2, J2)] if J2 else [])
    gm = q3.Model(L, L, 1, 1, q3.chain_bonds(L, terms))
    secs = []
    for N in range(1, 6):
        for K in range(L):
            s = q3.Sector.build(gm, N, (q3.wavevector(K, L), 0.0, 0.0))
            if s.dim:
                secs.append(s)
    kw = dict(maxitr=300, minitr=10, itrint=5, lnc_ene_conv=1e-12, prec=prec, seed=7, use_graph=graph)
    E1, it1, st1 = q3.lanczos_group(secs, **kw)
    q3.set_option("group_hop_cache", 0)
    try:
        E0, it0, st0 = q3.lanczos_group(secs, **kw)
    finally:
        q3.set_option("group_hop_cache", 1)
    assert st1["hop_cache_bytes"] > 0 and st0["hop_cache_bytes"] == 0
    assert np.array_equal(it1, it0)
    assert np.array_equal(E1, E0), np.max(np.abs(E1 - E0))


def test_lanczos_group_long_chain():
    """A group of long-chain sectors (L = 6000, N_down = 2, k = 0..3; about
    96 KB of LDS tables, beyond the default 64 KB): each E0 equals the
    per-sector plain Lanczos within 1e-9 (FP64)."""
    print("[gpu] lanczos group long chain", flush=True)
    q3 = _q3()
    L = 6000
    gm = q3.Model(L, L, 1, 1, orc.chain_bonds(L, [(1, 1.0, 0.8)]))
    secs = [q3.Sector.build(gm, 2, (q3.wavevector(K, L), 0.0, 0.0)) for K in range(4)]
    E, its, st = q3.lanczos_group(secs, maxitr=400, minitr=10, itrint=5, lnc_ene_conv=1e-12, prec=q3.FP64, seed=5,
                                  use_graph=True)
    for s, e in zip(secs, E):
        e1, _, _ = q3.lanczos_routines(s, maxitr=400, minitr=10, itrint=5, lnc_ene_conv=1e-12, prec=q3.FP64, seed=5,
                                       want_vector=False)
        assert abs(e - e1) <= 1e-9 * max(1.0, abs(e1)), (e, e1)


@pytest.mark.parametrize("L,steps", [(40, 30), (120, 20), (300, 12)])
@pytest.mark.parametrize("prec_name", ["fp64", "fp32"])
def test_block_stencil_matches_per_row_block_kernel(L, steps, prec_name):
    """The block gap-lattice stencil (clear rows of the N_down = 4 all-k
    sector) plus the per-row block kernel (the other rows) against the
    per-row block kernel on every row: same recurrence
    (get_expectation_values.f90:481-592) for all 64 q.  FP64 coefficients to
    1e-11 over the first steps; FP32 to the FP32 bar of this file."""
    print(f"[gpu] block stencil vs per-row block kernel L={L} {prec_name}", flush=True)
    q3 = _q3()
    prec = q3.FP64 if prec_name == "fp64" else q3.FP32
    gm = q3.Model(L, L, 1, 1, q3.chain_bonds(L, [(1, 1.0, 0.8)]))
    gs = q3.Sector.build(gm, 3, (0.0, 0.0, 0.0))
    E0, psi, _ = q3.lanczos_routines(gs, maxitr=2000, lnc_ene_conv=1e-13, prec=q3.FP64, seed=5)
    pos = np.array([[i, 0.0, 0.0] for i in range(L)], dtype=float)
    Ks = [int(round(L * m / 64)) % L for m in range(64)]
    try:
        q3.set_option("block_stencil", 0)
        a0, b0, n0, st0 = q3.dsf_sminus_batch(gs, Ks, pos, psi, steps, prec=prec)
    finally:
        q3.set_option("block_stencil", 1)
    q3.set_option("block_fused", 0)  # the separate H.v and update
    try:
        a1, b1, n1, st1 = q3.dsf_sminus_batch(gs, Ks, pos, psi, steps, prec=prec)
        # a second tile shape: 6 lines per tile, 256 threads, 8 slabs per task
        try:
            q3.set_option("block_stencil_lines", 6)
            q3.set_option("block_stencil_threads", 256)
            q3.set_option("block_stencil_slabs", 8)
            a2, b2, n2, st2 = q3.dsf_sminus_batch(gs, Ks, pos, psi, steps, prec=prec)
        finally:
            for o in ("block_stencil_lines", "block_stencil_threads", "block_stencil_slabs"):
                q3.set_option(o, 0)
    finally:
        q3.set_option("block_fused", 1)
    assert st0["block_stencil"] == 0 and st1["block_stencil"] == 1 and st2["block_stencil"] == 1
    assert 0 < st1["rows_per_row_kernel"] < st1["dim_all_k"]
    assert np.array_equal(n0, n1)
    k = 6 if prec == q3.FP64 else 4
    tol_a, tol_b = (1e-11, 1e-11) if prec == q3.FP64 else (1e-6, 1e-4)
    for j in range(64):
        if n0[j] <= 1e-20:
            continue
        for a_, b_ in ((a1, b1), (a2, b2)):
            ea = np.max(np.abs(a_[j, :k] - a0[j, :k])) / max(1.0, np.max(np.abs(a0[j, :k])))
            eb = np.max(np.abs(b_[j, :k] - b0[j, :k])) / max(1e-300, np.max(np.abs(b0[j, :k])))
            assert ea <= tol_a and eb <= tol_b, (j, Ks[j], ea, eb)


def test_block_stencil_autotune():
    """Option block_autotune: the batch times every block stencil tile shape
    on its S^- vectors, keeps the fastest (a later shape must win by 2 %),
    and its coefficients equal those of the default shape (FP64, 1e-11 over
    the first steps)."""
    print("[gpu] block stencil autotune", flush=True)
    q3 = _q3()
    L, steps = 120, 12
    gm = q3.Model(L, L, 1, 1, q3.chain_bonds(L, [(1, 1.0, 0.8)]))
    gs = q3.Sector.build(gm, 3, (0.0, 0.0, 0.0))
    E0, psi, _ = q3.lanczos_routines(gs, maxitr=2000, lnc_ene_conv=1e-13, prec=q3.FP64, seed=5)
    pos = np.array([[i, 0.0, 0.0] for i in range(L)], dtype=float)
    Ks = [int(round(L * m / 64)) % L for m in range(64)]
    a0, b0, n0, _ = q3.dsf_sminus_batch(gs, Ks, pos, psi, steps, prec=q3.FP64)
    assert len(q3.block_tune_last()) == 0
    q3.set_option("block_autotune", 1)
    try:
        a1, b1, n1, st1 = q3.dsf_sminus_batch(gs, Ks, pos, psi, steps, prec=q3.FP64)
    finally:
        q3.set_option("block_autotune", 0)
    tab = q3.block_tune_last()
    assert st1["block_stencil"] >= 1 and len(tab) >= 3
    timed, chosen = tab[:-1], tab[-1]
    assert all(r[3] > 0 for r in timed)
    assert chosen[3] <= min(r[3] for r in timed) / 0.98
    assert any(np.array_equal(chosen[:3], r[:3]) for r in timed)
    assert np.array_equal(n0, n1)
    for j in range(64):
        if n0[j] <= 1e-20:
            continue
        assert np.max(np.abs(a1[j, :6] - a0[j, :6])) <= 1e-11 * max(1.0, np.max(np.abs(a0[j, :6])))
        assert np.max(np.abs(b1[j, :6] - b0[j, :6])) <= 1e-11 * max(1e-300, np.max(np.abs(b0[j, :6])))


@pytest.mark.parametrize("L,steps", [(40, 30), (120, 20), (300, 12)])
@pytest.mark.parametrize("prec_name", ["fp64", "fp32"])
def test_block_fused_step_matches_separate_kernels(L, steps, prec_name):
    """The fused block step (option block_fused 1, the default on stencil
    sectors: the block stencil and the per-row block kernel form v_i from
    P_{i-1} and v_{i-1} as they read them and write v_i and P_i; alpha and
    beta per q from four sums) against the block stencil H.v followed by the
    block update (option block_fused 0), the recurrence of
    get_expectation_values.f90:481-592 for all 64 q.  FP64 coefficients to
    1e-10 over the first steps, FP32 to the FP32 bar of this file; also with a
    second tile shape (6 lines, 256 threads, 8 slabs per task)."""
    print(f"[gpu] fused block step vs separate kernels L={L} {prec_name}", flush=True)
    q3 = _q3()
    prec = q3.FP64 if prec_name == "fp64" else q3.FP32
    gm = q3.Model(L, L, 1, 1, q3.chain_bonds(L, [(1, 1.0, 0.8)]))
    gs = q3.Sector.build(gm, 3, (0.0, 0.0, 0.0))
    E0, psi, _ = q3.lanczos_routines(gs, maxitr=2000, lnc_ene_conv=1e-13, prec=q3.FP64, seed=5)
    pos = np.array([[i, 0.0, 0.0] for i in range(L)], dtype=float)
    Ks = [int(round(L * m / 64)) % L for m in range(64)]
    q3.set_option("block_fused", 0)
    try:
        a0, b0, n0, st0 = q3.dsf_sminus_batch(gs, Ks, pos, psi, steps, prec=prec)
    finally:
        q3.set_option("block_fused", 1)
    a1, b1, n1, st1 = q3.dsf_sminus_batch(gs, Ks, pos, psi, steps, prec=prec)
    try:
        q3.set_option("block_stencil_lines", 6)
        q3.set_option("block_stencil_threads", 256)
        q3.set_option("block_stencil_slabs", 8)
        a2, b2, n2, st2 = q3.dsf_sminus_batch(gs, Ks, pos, psi, steps, prec=prec)
    finally:
        for o in ("block_stencil_lines", "block_stencil_threads", "block_stencil_slabs"):
            q3.set_option(o, 0)
    assert st0["block_stencil"] == 1 and st1["block_stencil"] == 2 and st2["block_stencil"] == 2
    assert np.array_equal(n0, n1) and np.array_equal(n0, n2)
    k = 6 if prec == q3.FP64 else 4
    tol_a, tol_b = (1e-10, 1e-10) if prec == q3.FP64 else (1e-6, 1e-4)
    worst = [0.0, 0.0]
    for j in range(64):
        if n0[j] <= 1e-20:
            continue
        for a_, b_ in ((a1, b1), (a2, b2)):
            ea = np.max(np.abs(a_[j, :k] - a0[j, :k])) / max(1.0, np.max(np.abs(a0[j, :k])))
            eb = np.max(np.abs(b_[j, :k] - b0[j, :k])) / max(1e-300, np.max(np.abs(b0[j, :k])))
            worst = [max(worst[0], ea), max(worst[1], eb)]
            assert ea <= tol_a and eb <= tol_b, (j, Ks[j], ea, eb)
    print(f"[gpu]   worst alpha {worst[0]:.2e} beta^2 {worst[1]:.2e}", flush=True)
